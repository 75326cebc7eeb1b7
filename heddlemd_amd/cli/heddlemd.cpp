// heddlemd.cpp -- the `heddlemd run|lint` front end over libheddle.
//
// Reference front end: cli_main / cli_main_run / cli_main_lint
// (src/runner.rs:3292-3431), the TOML schema of src/io/config.rs
// (rqm/io/config-schema.md), the extended-XYZ init reader
// (src/io/init_state.rs) and trajectory writer (src/io/trajectory.rs:66-177),
// the CSV log (src/io/log_output.rs:74-103), the .minlog
// (src/io/minlog_output.rs) and .timings tables (src/timings.rs:520-550),
// the topology file (src/forces/topology.rs), and the Maxwell-Boltzmann
// start (src/runner.rs:3194-3283).  Usage, exit codes and output formats
// follow book/src/reference/cli.md.
//
// Each phase runs on a fresh libheddle context configured for the phase's
// slots; the particle state (positions, velocities, images, box) carries
// over between phases through the C ABI.
#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/heddle.h"
#include "toml.h"

namespace {

// ---- units (CODATA 2018, the reference's src/units/atomic_constants.rs) --
const double BOHR_M = 5.29177210903e-11;
const double ME_KG = 9.1093837015e-31;
const double E_C = 1.602176634e-19;
const double HARTREE_J = 4.359744722207101e-18;
const double AU_TIME_S = 2.4188843265857195e-17;
const double HARTREE_K = 315775.0248040668;  // E_h / k_B
const double AU_PRESSURE_PA = HARTREE_J / (BOHR_M * BOHR_M * BOHR_M);
const double AU_FORCE_N = HARTREE_J / BOHR_M;
const double AU_VELOCITY_MS = BOHR_M / AU_TIME_S;

enum Dim { LENGTH, MASS, CHARGE, ENERGY, TIME, TEMPERATURE, PRESSURE, FORCE, VELOCITY,
           INV_LENGTH, INV_TIME, INV_PRESSURE, STIFFNESS, NONE };

// SI value of one atomic unit of the dimension
double si_per_au(Dim d) {
  switch (d) {
    case LENGTH: return BOHR_M;
    case MASS: return ME_KG;
    case CHARGE: return E_C;
    case ENERGY: return HARTREE_J;
    case TIME: return AU_TIME_S;
    case TEMPERATURE: return HARTREE_K;
    case PRESSURE: return AU_PRESSURE_PA;
    case FORCE: return AU_FORCE_N;
    case VELOCITY: return AU_VELOCITY_MS;
    case INV_LENGTH: return 1.0 / BOHR_M;
    case INV_TIME: return 1.0 / AU_TIME_S;
    case INV_PRESSURE: return 1.0 / AU_PRESSURE_PA;
    case STIFFNESS: return HARTREE_J / (BOHR_M * BOHR_M);
    default: return 1.0;
  }
}

struct Units {
  bool si = true;
  double to_au(Dim d, double v) const { return si ? v / si_per_au(d) : v; }
  double to_user(Dim d, double v) const { return si ? v * si_per_au(d) : v; }
  double factor(Dim d) const { return si ? si_per_au(d) : 1.0; }
};

// ---- errors ------------------------------------------------------------
struct Failure {
  int code;  // exit code: 1 before the integration loop, 2 during
  std::string msg;
};
[[noreturn]] void fail1(const std::string& m) { throw Failure{1, m}; }
[[noreturn]] void fail2(const std::string& m) { throw Failure{2, m}; }

// ---- number formatting: Rust's `{:.Ne}` (no '+', no exponent padding) --
std::string fmt_e(double v, int digits = 9) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  char buf[64];
  snprintf(buf, sizeof buf, "%.*e", digits, v);
  std::string s(buf);
  size_t e = s.find('e');
  std::string mant = s.substr(0, e), ex = s.substr(e + 1);
  bool neg = ex[0] == '-';
  size_t k = (ex[0] == '-' || ex[0] == '+') ? 1 : 0;
  while (k + 1 < ex.size() && ex[k] == '0') ++k;
  return mant + "e" + (neg ? "-" : "") + ex.substr(k);
}

bool file_exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

std::string dir_of(const std::string& p) {
  size_t k = p.find_last_of('/');
  return k == std::string::npos ? "." : (k == 0 ? "/" : p.substr(0, k));
}
std::string base_of(const std::string& p) {
  size_t k = p.find_last_of('/');
  return k == std::string::npos ? p : p.substr(k + 1);
}
std::string resolve(const std::string& base, const std::string& p) {
  if (!p.empty() && p[0] == '/') return p;
  return base + "/" + p;
}

// ---- TOML access helpers ------------------------------------------------
const toml::Value* req(const toml::Value* t, const std::string& key, const std::string& where) {
  const toml::Value* v = t ? t->get(key) : nullptr;
  if (!v) fail1("missing field `" + (where.empty() ? key : where + "." + key) + "`");
  return v;
}
double num_of(const toml::Value* v, const std::string& name) {
  if (!v->is_number()) fail1("field `" + name + "` must be a number");
  return v->num();
}
double get_num(const toml::Value* t, const std::string& key, const std::string& where,
               const double* dflt = nullptr) {
  const toml::Value* v = t ? t->get(key) : nullptr;
  if (!v) {
    if (dflt) return *dflt;
    req(t, key, where);
  }
  return num_of(v, where + "." + key);
}
uint64_t get_u64(const toml::Value* t, const std::string& key, const std::string& where,
                 const uint64_t* dflt = nullptr) {
  const toml::Value* v = t ? t->get(key) : nullptr;
  if (!v) {
    if (dflt) return *dflt;
    req(t, key, where);
  }
  if (v->kind != toml::Value::INT || v->i < 0)
    fail1("field `" + where + "." + key + "` must be a non-negative integer");
  return (uint64_t)v->i;
}
std::string get_str(const toml::Value* t, const std::string& key, const std::string& where) {
  const toml::Value* v = req(t, key, where);
  if (v->kind != toml::Value::STRING) fail1("field `" + where + "." + key + "` must be a string");
  return v->s;
}
bool get_bool(const toml::Value* t, const std::string& key, bool dflt) {
  const toml::Value* v = t ? t->get(key) : nullptr;
  if (!v) return dflt;
  if (v->kind != toml::Value::BOOL) fail1("field `" + key + "` must be true or false");
  return v->b;
}
void check_keys(const toml::Value* t, const std::set<std::string>& allowed, const std::string& where) {
  if (!t) return;
  for (auto& kv : t->tab)
    if (!allowed.count(kv.first))
      fail1("unknown field `" + kv.first + "` in " + (where.empty() ? "the config" : "`" + where + "`"));
}

// ---- the configuration --------------------------------------------------
struct Slot {
  std::string kind;
  const toml::Value* params = nullptr;
};

struct PhaseCfg {
  bool minimize = false;
  std::string name;
  size_t order = 0;
  // MD
  uint64_t n_steps = 0;
  double dt = 0;
  Slot integrator, thermostat, barostat;
  std::string traj_path, log_path, timings_path, minlog_path;
  uint64_t traj_every = 0, log_every = 0, minlog_every = 1;
  bool include_velocities = true, include_images = true;
  // steepest descent (atomic units)
  double sd_initial = 0, sd_max = 0, sd_inc = 1.2, sd_dec = 0.2, sd_ftol = 0, sd_etol = 1e-7;
  uint64_t sd_max_iter = 1000;
};

struct ConstraintType {
  std::string name, kind;
  double d_oh = 0, d_hh = 0;  // settle
  unsigned atoms = 0;         // shake
  std::vector<std::tuple<unsigned, unsigned, double>> cons;
};

struct Config {
  std::string path, root, base_dir;
  Units units;
  std::string init_path, topology_path;
  uint64_t seed = 0;
  double temperature = 0;
  std::vector<std::string> type_names;
  std::vector<double> type_mass, type_charge;
  // LJ per type pair (ntypes x ntypes)
  std::vector<double> sigma, eps, cutoff, rswitch;
  std::vector<bool> pair_declared;
  bool have_spme = false;
  double alpha = 0, r_cut_real = 0;
  int grid[3] = {0, 0, 0};
  int spline_order = 4;
  bool all_pairs = false;
  double r_skin = -1;
  std::map<std::string, std::pair<int, std::vector<double>>> bond_types;      // kind 0 harmonic (k, r0), 1 morse (de, a, re)
  std::map<std::string, std::vector<double>> angle_types;                     // k, theta0
  std::map<std::string, std::vector<double>> dihedral_types;                  // k, n, phi0, s_lj, s_coul
  std::vector<ConstraintType> ctypes;
  std::vector<PhaseCfg> phases;
  std::shared_ptr<toml::Value> doc;

  int ntypes() const { return (int)type_names.size(); }
  int type_index(const std::string& n) const {
    for (int k = 0; k < ntypes(); ++k)
      if (type_names[k] == n) return k;
    return -1;
  }
  double max_cutoff() const {
    double m = have_spme ? r_cut_real : 0.0;
    for (double c : cutoff) m = std::max(m, c);
    return m;
  }
};

std::string read_file(const std::string& p) {
  std::ifstream f(p, std::ios::binary);
  if (!f) fail1("cannot read `" + p + "`");
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

Slot get_slot(const toml::Value* t, const std::string& key, const std::string& where, bool required) {
  const toml::Value* v = t->get(key);
  if (!v) {
    if (required) req(t, key, where);
    return Slot{};
  }
  return Slot{get_str(v, "kind", where + "." + key), v};
}

Config load_config(const std::string& path) {
  Config c;
  c.path = path;
  std::string fname = base_of(path);
  const std::string suf = ".in.toml";
  if (fname.size() <= suf.size() || fname.compare(fname.size() - suf.size(), suf.size(), suf))
    fail1("config file name must end in `.in.toml`: `" + path + "`");
  c.root = fname.substr(0, fname.size() - suf.size());
  c.base_dir = dir_of(path);
  std::string text = read_file(path);
  try {
    c.doc = toml::parse(text);
  } catch (const toml::ParseError& e) {
    fail1(path + ": " + e.what());
  }
  const toml::Value* d = c.doc.get();
  check_keys(d, {"schema_version", "units", "init", "topology", "simulation", "phase",
                 "minimization", "particle_types", "pair_interactions", "bond_types",
                 "angle_types", "dihedral_types", "constraint_types", "spme", "neighbor_list"},
             "");
  if (get_u64(d, "schema_version", "") != 1) fail1("unsupported schema_version (1 is supported)");
  if (const toml::Value* u = d->get("units")) {
    if (u->kind != toml::Value::STRING || (u->s != "si" && u->s != "atomic"))
      fail1("unknown units `" + u->s + "` (si or atomic)");
    c.units.si = u->s == "si";
  }
  const Units& U = c.units;
  c.init_path = resolve(c.base_dir, get_str(d, "init", ""));
  if (d->get("topology")) c.topology_path = resolve(c.base_dir, get_str(d, "topology", ""));
  const toml::Value* sim = req(d, "simulation", "");
  check_keys(sim, {"seed", "temperature", "graph_batch_size", "cuda_graphs_disable", "fast_math"},
             "simulation");
  c.seed = get_u64(sim, "seed", "simulation");
  c.temperature = U.to_au(TEMPERATURE, get_num(sim, "temperature", "simulation"));

  const toml::Value* pts = req(d, "particle_types", "");
  if (pts->kind != toml::Value::ARRAY || pts->arr.empty()) fail1("`particle_types` must list at least one type");
  for (auto& pt : pts->arr) {
    check_keys(pt.get(), {"name", "mass", "charge"}, "particle_types");
    std::string n = get_str(pt.get(), "name", "particle_types");
    if (c.type_index(n) >= 0) fail1("particle type `" + n + "` declared twice");
    double zero = 0.0;
    c.type_names.push_back(n);
    c.type_mass.push_back(U.to_au(MASS, get_num(pt.get(), "mass", "particle_types")));
    c.type_charge.push_back(U.to_au(CHARGE, get_num(pt.get(), "charge", "particle_types", &zero)));
    if (!(c.type_mass.back() > 0)) fail1("particle type `" + n + "`: mass must be > 0");
  }
  int nt = c.ntypes();
  c.sigma.assign(nt * nt, 1.0);
  c.eps.assign(nt * nt, 0.0);
  c.cutoff.assign(nt * nt, 0.0);
  c.rswitch.assign(nt * nt, 0.0);
  c.pair_declared.assign(nt * nt, false);
  if (const toml::Value* pis = d->get("pair_interactions")) {
    for (auto& pi : pis->arr) {
      check_keys(pi.get(), {"between", "potential", "sigma", "epsilon", "cutoff", "r_switch"},
                 "pair_interactions");
      if (get_str(pi.get(), "potential", "pair_interactions") != "lennard-jones")
        fail1("pair_interactions: only potential = \"lennard-jones\" is known");
      const toml::Value* bw = req(pi.get(), "between", "pair_interactions");
      if (bw->kind != toml::Value::ARRAY || bw->arr.size() != 2) fail1("`between` must name two types");
      int a = c.type_index(bw->arr[0]->s), b = c.type_index(bw->arr[1]->s);
      if (a < 0 || b < 0) fail1("pair_interactions: unknown particle type in `between`");
      double sg = U.to_au(LENGTH, get_num(pi.get(), "sigma", "pair_interactions"));
      double ep = U.to_au(ENERGY, get_num(pi.get(), "epsilon", "pair_interactions"));
      double rc = U.to_au(LENGTH, get_num(pi.get(), "cutoff", "pair_interactions"));
      double rs = pi->get("r_switch") ? U.to_au(LENGTH, get_num(pi.get(), "r_switch", "pair_interactions")) : rc;
      if (c.pair_declared[a * nt + b]) fail1("pair interaction declared twice");
      for (int k : {a * nt + b, b * nt + a}) {
        c.sigma[k] = sg; c.eps[k] = ep; c.cutoff[k] = rc; c.rswitch[k] = rs;
        c.pair_declared[k] = true;
      }
    }
  }
  for (int k = 0; k < nt * nt; ++k)
    if (!c.pair_declared[k])
      fail1("pair interaction between `" + c.type_names[k / nt] + "` and `" +
            c.type_names[k % nt] + "` is not declared");
  if (const toml::Value* sp = d->get("spme")) {
    check_keys(sp, {"alpha", "r_cut_real", "grid", "spline_order"}, "spme");
    c.have_spme = true;
    c.alpha = U.to_au(INV_LENGTH, get_num(sp, "alpha", "spme"));
    c.r_cut_real = U.to_au(LENGTH, get_num(sp, "r_cut_real", "spme"));
    const toml::Value* g = req(sp, "grid", "spme");
    if (g->kind != toml::Value::ARRAY || g->arr.size() != 3) fail1("`spme.grid` must have three entries");
    for (int k = 0; k < 3; ++k) c.grid[k] = (int)g->arr[k]->i;
    uint64_t four = 4;
    c.spline_order = (int)get_u64(sp, "spline_order", "spme", &four);
  }
  if (const toml::Value* nl = d->get("neighbor_list")) {
    std::string mode = get_str(nl, "mode", "neighbor_list");
    if (mode == "all-pairs") {
      check_keys(nl, {"mode"}, "neighbor_list");
      c.all_pairs = true;
    } else if (mode == "cell-list") {
      check_keys(nl, {"mode", "r_skin"}, "neighbor_list");
      if (nl->get("r_skin")) c.r_skin = U.to_au(LENGTH, get_num(nl, "r_skin", "neighbor_list"));
    } else {
      fail1("neighbor_list.mode must be \"cell-list\" or \"all-pairs\"");
    }
  }
  if (const toml::Value* bts = d->get("bond_types"))
    for (auto& bt : bts->arr) {
      std::string n = get_str(bt.get(), "name", "bond_types");
      std::string pot = get_str(bt.get(), "potential", "bond_types");
      if (pot == "harmonic") {
        check_keys(bt.get(), {"name", "potential", "k", "r0"}, "bond_types");
        c.bond_types[n] = {0, {U.to_au(STIFFNESS, get_num(bt.get(), "k", "bond_types")),
                               U.to_au(LENGTH, get_num(bt.get(), "r0", "bond_types"))}};
      } else if (pot == "morse") {
        check_keys(bt.get(), {"name", "potential", "de", "a", "re"}, "bond_types");
        c.bond_types[n] = {1, {U.to_au(ENERGY, get_num(bt.get(), "de", "bond_types")),
                               U.to_au(INV_LENGTH, get_num(bt.get(), "a", "bond_types")),
                               U.to_au(LENGTH, get_num(bt.get(), "re", "bond_types"))}};
      } else {
        fail1("bond_types: unknown potential `" + pot + "`");
      }
    }
  if (const toml::Value* ats = d->get("angle_types"))
    for (auto& at : ats->arr) {
      check_keys(at.get(), {"name", "potential", "k_theta", "theta_0"}, "angle_types");
      if (get_str(at.get(), "potential", "angle_types") != "harmonic")
        fail1("angle_types: only potential = \"harmonic\" is known");
      c.angle_types[get_str(at.get(), "name", "angle_types")] = {
          U.to_au(ENERGY, get_num(at.get(), "k_theta", "angle_types")),
          get_num(at.get(), "theta_0", "angle_types")};
    }
  if (const toml::Value* dts = d->get("dihedral_types"))
    for (auto& dt : dts->arr) {
      check_keys(dt.get(), {"name", "potential", "k_phi", "n", "phi_0", "scale_lj_14", "scale_coul_14"},
                 "dihedral_types");
      if (get_str(dt.get(), "potential", "dihedral_types") != "periodic")
        fail1("dihedral_types: only potential = \"periodic\" is known");
      double slj = 0.5, scoul = 1.0 / 1.2;  // AMBER defaults
      c.dihedral_types[get_str(dt.get(), "name", "dihedral_types")] = {
          U.to_au(ENERGY, get_num(dt.get(), "k_phi", "dihedral_types")),
          (double)get_u64(dt.get(), "n", "dihedral_types"), get_num(dt.get(), "phi_0", "dihedral_types"),
          get_num(dt.get(), "scale_lj_14", "dihedral_types", &slj),
          get_num(dt.get(), "scale_coul_14", "dihedral_types", &scoul)};
    }
  if (const toml::Value* cts = d->get("constraint_types"))
    for (auto& ct : cts->arr) {
      ConstraintType t;
      t.name = get_str(ct.get(), "name", "constraint_types");
      t.kind = get_str(ct.get(), "kind", "constraint_types");
      if (t.kind == "settle") {
        check_keys(ct.get(), {"name", "kind", "d_OH", "d_HH"}, "constraint_types");
        t.d_oh = U.to_au(LENGTH, get_num(ct.get(), "d_OH", "constraint_types"));
        t.d_hh = U.to_au(LENGTH, get_num(ct.get(), "d_HH", "constraint_types"));
        t.atoms = 3;
      } else if (t.kind == "shake") {
        check_keys(ct.get(), {"name", "kind", "atoms", "constraints"}, "constraint_types");
        t.atoms = (unsigned)get_u64(ct.get(), "atoms", "constraint_types");
        const toml::Value* cl = req(ct.get(), "constraints", "constraint_types");
        for (auto& e : cl->arr)
          t.cons.emplace_back((unsigned)get_u64(e.get(), "i", "constraints"),
                              (unsigned)get_u64(e.get(), "j", "constraints"),
                              U.to_au(LENGTH, get_num(e.get(), "d", "constraints")));
      } else {
        fail1("constraint_types: unknown kind `" + t.kind + "` (settle, shake)");
      }
      c.ctypes.push_back(t);
    }

  // phases and minimizations in document order
  auto out_path = [&](const toml::Value* o, const char* key, const std::string& name,
                      const char* ext) {
    if (o && o->get(key)) return resolve(c.base_dir, get_str(o, key, "output"));
    return c.base_dir + "/" + c.root + ".out." + name + "." + ext;
  };
  if (const toml::Value* ps = d->get("phase"))
    for (auto& pv : ps->arr) {
      const toml::Value* p = pv.get();
      check_keys(p, {"name", "n_steps", "dt", "integrator", "thermostat", "barostat", "output"}, "phase");
      PhaseCfg ph;
      ph.name = get_str(p, "name", "phase");
      ph.order = p->order;
      ph.n_steps = get_u64(p, "n_steps", "phase");
      ph.dt = U.to_au(TIME, get_num(p, "dt", "phase"));
      if (!(ph.dt > 0)) fail1("phase `" + ph.name + "`: dt must be > 0");
      ph.integrator = get_slot(p, "integrator", "phase", true);
      ph.thermostat = get_slot(p, "thermostat", "phase", false);
      ph.barostat = get_slot(p, "barostat", "phase", false);
      const toml::Value* o = p->get("output");
      check_keys(o, {"trajectory_path", "trajectory_every", "include_velocities", "include_images",
                     "log_path", "log_every", "timings_path"},
                 "phase.output");
      uint64_t tr = 100, lg = 100;  // config.rs:595-600
      ph.traj_every = get_u64(o, "trajectory_every", "phase.output", &tr);
      ph.log_every = get_u64(o, "log_every", "phase.output", &lg);
      ph.include_velocities = get_bool(o, "include_velocities", true);
      ph.include_images = get_bool(o, "include_images", true);
      ph.traj_path = out_path(o, "trajectory_path", ph.name, "xyz");
      ph.log_path = out_path(o, "log_path", ph.name, "log");
      ph.timings_path = out_path(o, "timings_path", ph.name, "timings");
      c.phases.push_back(ph);
    }
  if (const toml::Value* ms = d->get("minimization"))
    for (auto& mv : ms->arr) {
      const toml::Value* m = mv.get();
      check_keys(m, {"name", "algorithm", "output"}, "minimization");
      PhaseCfg ph;
      ph.minimize = true;
      ph.name = get_str(m, "name", "minimization");
      ph.order = m->order;
      Slot alg = get_slot(m, "algorithm", "minimization", true);
      if (alg.kind != "steepest-descent")
        fail1("minimization.algorithm: unknown kind `" + alg.kind + "`");
      const toml::Value* a = alg.params;
      check_keys(a, {"kind", "initial_step", "max_step", "step_increase", "step_decrease",
                     "force_tolerance", "energy_tolerance", "max_iterations"},
                 "minimization.algorithm");
      double d_init = 1.0e-12, d_max = 1.0e-10, d_inc = 1.2, d_dec = 0.2, d_ftol = 1.0e-10,
             d_etol = 1.0e-7;
      uint64_t d_iter = 1000;
      ph.sd_initial = U.to_au(LENGTH, get_num(a, "initial_step", "minimization.algorithm", &d_init));
      ph.sd_max = U.to_au(LENGTH, get_num(a, "max_step", "minimization.algorithm", &d_max));
      ph.sd_inc = get_num(a, "step_increase", "minimization.algorithm", &d_inc);
      ph.sd_dec = get_num(a, "step_decrease", "minimization.algorithm", &d_dec);
      ph.sd_ftol = U.to_au(FORCE, get_num(a, "force_tolerance", "minimization.algorithm", &d_ftol));
      ph.sd_etol = get_num(a, "energy_tolerance", "minimization.algorithm", &d_etol);
      ph.sd_max_iter = get_u64(a, "max_iterations", "minimization.algorithm", &d_iter);
      const toml::Value* o = m->get("output");
      check_keys(o, {"minlog_path", "minlog_every", "trajectory_path", "trajectory_every",
                     "include_images", "timings_path"},
                 "minimization.output");
      uint64_t one = 1, zero = 0;
      ph.minlog_every = get_u64(o, "minlog_every", "minimization.output", &one);
      ph.traj_every = get_u64(o, "trajectory_every", "minimization.output", &zero);
      ph.include_velocities = false;
      ph.include_images = get_bool(o, "include_images", true);
      ph.minlog_path = out_path(o, "minlog_path", ph.name, "minlog");
      ph.traj_path = out_path(o, "trajectory_path", ph.name, "xyz");
      ph.timings_path = out_path(o, "timings_path", ph.name, "timings");
      c.phases.push_back(ph);
    }
  // registry kinds (src/registries.rs): the reference's kinds are known;
  // those without a device path here are refused by name
  for (auto& ph : c.phases) {
    if (ph.minimize) continue;
    auto check = [&](const Slot& sl, const char* slot, std::set<std::string> built,
                     std::set<std::string> known) {
      if (sl.kind.empty()) return;
      if (built.count(sl.kind)) return;
      if (known.count(sl.kind))
        fail1("phase `" + ph.name + "`: " + slot + " kind `" + sl.kind + "` is not built here");
      fail1("phase `" + ph.name + "`: unknown " + slot + " kind `" + sl.kind + "`");
    };
    check(ph.integrator, "integrator", {"velocity-verlet", "langevin-baoab", "mtk-npt"}, {});
    // mtk-npt owns the thermostat and the barostat (mtk_npt.rs:572-573)
    if (ph.integrator.kind == "mtk-npt" && !ph.thermostat.kind.empty())
      fail1("phase `" + ph.name + "`: integrator `mtk-npt` owns the thermostat; remove "
            "[phase.thermostat]");
    if (ph.integrator.kind == "mtk-npt" && !ph.barostat.kind.empty())
      fail1("phase `" + ph.name + "`: integrator `mtk-npt` owns the barostat; remove "
            "[phase.barostat]");
    check(ph.thermostat, "thermostat", {"nose-hoover-chain", "csvr", "berendsen", "andersen"}, {});
    check(ph.barostat, "barostat", {"c-rescale", "berendsen", "monte-carlo"}, {});
  }
  std::sort(c.phases.begin(), c.phases.end(),
            [](const PhaseCfg& a, const PhaseCfg& b) { return a.order < b.order; });
  std::set<std::string> names;
  for (auto& p : c.phases)
    if (!names.insert(p.name).second) fail1("phase name `" + p.name + "` used twice");
  return c;
}

// ---- particle state -----------------------------------------------------
struct State {
  uint64_t n = 0;
  std::vector<float> posq;  // x, y, z, q (atomic units)
  std::vector<float> mass;
  std::vector<double> mass_d;  // the type masses in f64
  std::vector<uint32_t> type;
  std::vector<float> vx, vy, vz;
  std::vector<int32_t> ix, iy, iz;
  double lattice[6] = {0, 0, 0, 0, 0, 0};
  bool have_velocities = false;
};

std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

double parse_double(const std::string& s, const std::string& ctx) {
  char* end = nullptr;
  double v = strtod(s.c_str(), &end);
  if (!end || *end) fail1(ctx + ": malformed number `" + s + "`");
  return v;
}

// Extended XYZ (src/io/init_state.rs): N, a comment line with
// Lattice="lx 0 0 xy ly 0 xz yz lz" and Properties=species:S:1:pos:R:3
// [:velo:R:3][:image:I:3], then one row per particle.
State load_init(const Config& c) {
  std::ifstream f(c.init_path);
  if (!f) fail1("cannot read init file `" + c.init_path + "`");
  std::string line;
  State s;
  if (!std::getline(f, line)) fail1(c.init_path + ": empty init file");
  s.n = (uint64_t)parse_double(split_ws(line).at(0), c.init_path + " line 1");
  if (!std::getline(f, line)) fail1(c.init_path + ": missing the comment line");
  size_t lp = line.find("Lattice=\"");
  if (lp == std::string::npos) fail1(c.init_path + ": the comment line has no Lattice=\"...\"");
  size_t le = line.find('"', lp + 9);
  std::vector<std::string> lv = split_ws(line.substr(lp + 9, le - lp - 9));
  if (lv.size() != 9) fail1(c.init_path + ": Lattice needs nine numbers");
  double m9[9];
  for (int k = 0; k < 9; ++k) m9[k] = c.units.to_au(LENGTH, parse_double(lv[k], "Lattice"));
  if (m9[1] != 0 || m9[2] != 0 || m9[5] != 0)
    fail1(c.init_path + ": the lattice must be lower triangular");
  s.lattice[0] = m9[0]; s.lattice[1] = m9[4]; s.lattice[2] = m9[8];
  s.lattice[3] = m9[3]; s.lattice[4] = m9[6]; s.lattice[5] = m9[7];
  size_t pp = line.find("Properties=");
  std::string props = pp == std::string::npos ? "species:S:1:pos:R:3"
                                              : split_ws(line.substr(pp + 11)).at(0);
  bool has_v = props.find("velo:R:3") != std::string::npos;
  bool has_i = props.find("image:I:3") != std::string::npos;
  if (props.compare(0, 19, "species:S:1:pos:R:3") != 0)
    fail1(c.init_path + ": Properties must start with species:S:1:pos:R:3");
  s.posq.resize(4 * s.n);
  s.mass.resize(s.n);
  s.mass_d.resize(s.n);
  s.type.resize(s.n);
  s.vx.assign(s.n, 0.f); s.vy.assign(s.n, 0.f); s.vz.assign(s.n, 0.f);
  s.ix.assign(s.n, 0); s.iy.assign(s.n, 0); s.iz.assign(s.n, 0);
  for (uint64_t i = 0; i < s.n; ++i) {
    if (!std::getline(f, line)) fail1(c.init_path + ": fewer particle rows than declared");
    auto tok = split_ws(line);
    size_t need = 4 + (has_v ? 3 : 0) + (has_i ? 3 : 0);
    if (tok.size() < need) fail1(c.init_path + ": row " + std::to_string(i + 3) + " is short");
    int t = c.type_index(tok[0]);
    if (t < 0) fail1(c.init_path + ": unknown species `" + tok[0] + "`");
    s.type[i] = (uint32_t)t;
    for (int d = 0; d < 3; ++d)
      s.posq[4 * i + d] = (float)c.units.to_au(LENGTH, parse_double(tok[1 + d], "position"));
    s.posq[4 * i + 3] = (float)c.type_charge[t];
    s.mass[i] = (float)c.type_mass[t];
    s.mass_d[i] = c.type_mass[t];
    size_t k = 4;
    if (has_v) {
      s.vx[i] = (float)c.units.to_au(VELOCITY, parse_double(tok[k], "velocity"));
      s.vy[i] = (float)c.units.to_au(VELOCITY, parse_double(tok[k + 1], "velocity"));
      s.vz[i] = (float)c.units.to_au(VELOCITY, parse_double(tok[k + 2], "velocity"));
      k += 3;
    }
    if (has_i) {
      s.ix[i] = (int32_t)parse_double(tok[k], "image");
      s.iy[i] = (int32_t)parse_double(tok[k + 1], "image");
      s.iz[i] = (int32_t)parse_double(tok[k + 2], "image");
    }
  }
  s.have_velocities = has_v;
  return s;
}

// ---- topology (src/forces/topology.rs) ------------------------------------
struct Topology {
  std::vector<uint32_t> bond_ij;
  std::vector<int> bond_kind;
  std::vector<double> bond_p;  // 3 per bond
  std::vector<uint32_t> angle_ijk;
  std::vector<double> angle_k, angle_t0;
  std::vector<uint32_t> dih_ijkl;
  std::vector<double> dih_k, dih_n, dih_phi0;
  std::map<std::pair<uint32_t, uint32_t>, std::pair<double, double>> excl;  // (lo, hi) -> scales
  std::vector<std::vector<uint32_t>> groups;
  std::vector<int> group_type;
  uint64_t n_constraints = 0;
  // molecules: connected components of bonds and constraint groups
  // (MoleculeList::from_topology, topology.rs:1496-1560), ordered by their
  // smallest atom, atoms ascending; empty when every atom stands alone
  std::vector<uint32_t> mol_off, mol_atoms;
};

Topology load_topology(const Config& c, State& st) {
  Topology t;
  if (c.topology_path.empty()) return t;
  std::ifstream f(c.topology_path);
  if (!f) fail1("cannot read topology file `" + c.topology_path + "`");
  std::string line, section;
  int lineno = 0;
  std::map<std::pair<uint32_t, uint32_t>, std::pair<double, double>> explicit_ex;
  auto atom = [&](const std::string& s) {
    double v = parse_double(s, c.topology_path + " line " + std::to_string(lineno));
    if (v < 0 || v >= (double)st.n) fail1(c.topology_path + " line " + std::to_string(lineno) + ": atom index out of range");
    return (uint32_t)v;
  };
  struct Dih { uint32_t a[4]; std::string type; };
  std::vector<Dih> dihs;
  while (std::getline(f, line)) {
    ++lineno;
    size_t h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    auto tok = split_ws(line);
    if (tok.empty()) continue;
    if (tok[0][0] == '[') {
      section = tok[0];
      continue;
    }
    std::string where = c.topology_path + " line " + std::to_string(lineno);
    if (section == "[bonds]") {
      if (tok.size() != 3) fail1(where + ": a bond row is `i j type`");
      auto it = c.bond_types.find(tok[2]);
      if (it == c.bond_types.end()) fail1(where + ": unknown bond type `" + tok[2] + "`");
      t.bond_ij.push_back(atom(tok[0]));
      t.bond_ij.push_back(atom(tok[1]));
      t.bond_kind.push_back(it->second.first);
      for (int k = 0; k < 3; ++k) t.bond_p.push_back(k < (int)it->second.second.size() ? it->second.second[k] : 0.0);
    } else if (section == "[angles]") {
      if (tok.size() != 4) fail1(where + ": an angle row is `i j k type`");
      auto it = c.angle_types.find(tok[3]);
      if (it == c.angle_types.end()) fail1(where + ": unknown angle type `" + tok[3] + "`");
      for (int k = 0; k < 3; ++k) t.angle_ijk.push_back(atom(tok[k]));
      t.angle_k.push_back(it->second[0]);
      t.angle_t0.push_back(it->second[1]);
    } else if (section == "[dihedrals]") {
      if (tok.size() != 5) fail1(where + ": a dihedral row is `i j k l type`");
      auto it = c.dihedral_types.find(tok[4]);
      if (it == c.dihedral_types.end()) fail1(where + ": unknown dihedral type `" + tok[4] + "`");
      Dih dd;
      for (int k = 0; k < 4; ++k) dd.a[k] = atom(tok[k]);
      dd.type = tok[4];
      dihs.push_back(dd);
      for (int k = 0; k < 4; ++k) t.dih_ijkl.push_back(dd.a[k]);
      t.dih_k.push_back(it->second[0]);
      t.dih_n.push_back(it->second[1]);
      t.dih_phi0.push_back(it->second[2]);
    } else if (section == "[exclusions]") {
      if (tok.size() != 4) fail1(where + ": an exclusion row is `i j scale_lj scale_coul`");
      uint32_t a = atom(tok[0]), b = atom(tok[1]);
      double slj = parse_double(tok[2], where), sc = parse_double(tok[3], where);
      if (a == b || slj < 0 || slj > 1 || sc < 0 || sc > 1) fail1(where + ": invalid exclusion row");
      auto key = std::make_pair(std::min(a, b), std::max(a, b));
      if (explicit_ex.count(key)) fail1(where + ": duplicate exclusion");
      explicit_ex[key] = {slj, sc};
    } else if (section == "[constraints]") {
      if (tok.size() < 3) fail1(where + ": a constraint row lists atoms then a type");
      int ct = -1;
      for (size_t k = 0; k < c.ctypes.size(); ++k)
        if (c.ctypes[k].name == tok.back()) ct = (int)k;
      if (ct < 0) fail1(where + ": unknown constraint type `" + tok.back() + "`");
      std::vector<uint32_t> g;
      for (size_t k = 0; k + 1 < tok.size(); ++k) g.push_back(atom(tok[k]));
      if (g.size() != c.ctypes[ct].atoms)
        fail1(where + ": constraint type `" + tok.back() + "` takes " +
              std::to_string(c.ctypes[ct].atoms) + " atoms");
      t.groups.push_back(g);
      t.group_type.push_back(ct);
      t.n_constraints += c.ctypes[ct].kind == "settle" ? 3 : c.ctypes[ct].cons.size();
    } else if (section == "[charges]") {
      if (tok.size() != 2) fail1(where + ": a charge row is `i q`");
      st.posq[4 * atom(tok[0]) + 3] = (float)c.units.to_au(CHARGE, parse_double(tok[1], where));
    } else {
      fail1(where + ": row outside a known section (" + section + ")");
    }
  }
  // effective exclusions: explicit rows, then bonds (1-2), angles (1-3)
  // and constraint groups (all pairs) fully excluded, then scaled 1-4
  // pairs of dihedrals; the first entry for a pair wins
  t.excl = explicit_ex;
  auto add = [&](uint32_t a, uint32_t b, double slj, double sc) {
    auto key = std::make_pair(std::min(a, b), std::max(a, b));
    if (!t.excl.count(key)) t.excl[key] = {slj, sc};
  };
  for (size_t k = 0; k < t.bond_ij.size(); k += 2) add(t.bond_ij[k], t.bond_ij[k + 1], 0, 0);
  for (size_t k = 0; k < t.angle_ijk.size(); k += 3) add(t.angle_ijk[k], t.angle_ijk[k + 2], 0, 0);
  for (auto& g : t.groups)
    for (size_t a = 0; a < g.size(); ++a)
      for (size_t b = a + 1; b < g.size(); ++b) add(g[a], g[b], 0, 0);
  for (auto& dd : dihs) {
    auto& ty = c.dihedral_types.at(dd.type);
    add(dd.a[0], dd.a[3], ty[3], ty[4]);
  }
  if (!t.bond_ij.empty() || !t.groups.empty()) {
    std::vector<uint32_t> parent(st.n);
    for (uint64_t i = 0; i < st.n; ++i) parent[i] = (uint32_t)i;
    std::function<uint32_t(uint32_t)> root = [&](uint32_t x) {
      while (parent[x] != x) x = parent[x] = parent[parent[x]];
      return x;
    };
    auto join = [&](uint32_t a, uint32_t b) {
      uint32_t ra = root(a), rb = root(b);
      if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
    };
    for (size_t k = 0; k < t.bond_ij.size(); k += 2) join(t.bond_ij[k], t.bond_ij[k + 1]);
    for (auto& g : t.groups)
      for (size_t k = 1; k < g.size(); ++k) join(g[0], g[k]);
    // roots are each component's smallest atom: bucket atoms by root
    std::vector<uint32_t> count(st.n, 0), start(st.n, 0);
    for (uint64_t i = 0; i < st.n; ++i) count[root((uint32_t)i)]++;
    t.mol_off.push_back(0);
    uint32_t acc = 0;
    for (uint64_t i = 0; i < st.n; ++i)
      if (count[i]) {
        start[i] = acc;
        acc += count[i];
        t.mol_off.push_back(acc);
      }
    t.mol_atoms.resize(st.n);
    for (uint64_t i = 0; i < st.n; ++i) t.mol_atoms[start[root((uint32_t)i)]++] = (uint32_t)i;
  }
  return t;
}

// ---- Maxwell-Boltzmann start (src/runner.rs:3194-3283) --------------------
// rand_chacha 0.3 ChaCha8Rng::seed_from_u64: PCG32 key expansion, ChaCha
// with 8 rounds, 64-bit block counter, u64 = two LE words; rand 0.8
// gen::<f64>() = (u64 >> 11) * 2^-53.
struct ChaCha8 {
  uint32_t key[8];
  uint64_t block = 0;
  uint32_t buf[16];
  int used = 16;
  explicit ChaCha8(uint64_t seed) {
    uint64_t state = seed;
    const uint64_t mul = 6364136223846793005ull, inc = 11634580027462260723ull;
    for (int k = 0; k < 8; ++k) {
      state = state * mul + inc;
      uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
      uint32_t rot = (uint32_t)(state >> 59);
      key[k] = (xs >> rot) | (xs << ((32 - rot) & 31));
    }
  }
  static uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
  void refill() {
    uint32_t st[16] = {0x61707865, 0x3320646E, 0x79622D32, 0x6B206574};
    for (int k = 0; k < 8; ++k) st[4 + k] = key[k];
    st[12] = (uint32_t)block;
    st[13] = (uint32_t)(block >> 32);
    st[14] = st[15] = 0;
    uint32_t x[16];
    memcpy(x, st, sizeof x);
    auto qr = [&](int a, int b, int c2, int d) {
      x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16);
      x[c2] += x[d]; x[b] = rotl(x[b] ^ x[c2], 12);
      x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);
      x[c2] += x[d]; x[b] = rotl(x[b] ^ x[c2], 7);
    };
    for (int r = 0; r < 4; ++r) {
      qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
      qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
    }
    for (int k = 0; k < 16; ++k) buf[k] = x[k] + st[k];
    ++block;
    used = 0;
  }
  uint64_t next_u64() {
    if (used > 14) refill();
    uint64_t lo = buf[used], hi = buf[used + 1];
    used += 2;
    return lo | (hi << 32);
  }
  double next_f64() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
};

void maxwell_boltzmann(State& s, const std::vector<double>& mass, double kt, uint64_t seed,
                       uint64_t n_constraints) {
  uint64_t n = s.n;
  if (kt == 0.0 || n == 0) return;
  ChaCha8 rng(seed);
  for (uint64_t i = 0; i < n; ++i) {
    double sig = sqrt(kt / mass[i]);
    float* out[3] = {&s.vx[i], &s.vy[i], &s.vz[i]};
    for (int d = 0; d < 3; ++d) {
      double u1 = 1.0 - rng.next_f64(), u2 = rng.next_f64();
      double z = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
      *out[d] = (float)(z * sig);
    }
  }
  // f64 type masses and sums in index order (runner.rs:3224-3242)
  double total = 0;
  for (uint64_t i = 0; i < n; ++i) total += mass[i];
  std::vector<float>* comp[3] = {&s.vx, &s.vy, &s.vz};
  if (total > 0)
    for (int d = 0; d < 3; ++d) {
      double p = 0;
      for (uint64_t i = 0; i < n; ++i) p += mass[i] * (double)(*comp[d])[i];
      for (uint64_t i = 0; i < n; ++i) (*comp[d])[i] = (float)((double)(*comp[d])[i] - p / total);
    }
  long long ndof = 3LL * (long long)n - (long long)n_constraints - 3;
  if (ndof <= 0) {
    for (int d = 0; d < 3; ++d) std::fill(comp[d]->begin(), comp[d]->end(), 0.f);
    return;
  }
  double ke = 0;
  for (uint64_t i = 0; i < n; ++i) {
    double x = s.vx[i], y = s.vy[i], z = s.vz[i];
    ke += 0.5 * mass[i] * (x * x + y * y + z * z);
  }
  if (ke > 0) {
    double scale = sqrt(0.5 * (double)ndof * kt / ke);
    for (int d = 0; d < 3; ++d)
      for (uint64_t i = 0; i < n; ++i) (*comp[d])[i] = (float)((double)(*comp[d])[i] * scale);
  }
}

// ---- writers -------------------------------------------------------------
struct OutFile {
  FILE* f = nullptr;
  std::string path;
  void open(const std::string& p) {
    path = p;
    f = fopen(p.c_str(), "wx");  // fails when the file exists
    if (!f) fail1("output file already exists or cannot be created: `" + p + "`");
  }
  void put(const std::string& s) {
    if (fwrite(s.data(), 1, s.size(), f) != s.size()) fail2("failed to write `" + path + "`");
  }
  void close() {
    if (f && fclose(f) != 0) fail2("failed to write `" + path + "`");
    f = nullptr;
  }
  ~OutFile() {
    if (f) fclose(f);
  }
};

void write_frame(OutFile& o, const Config& c, uint64_t step, double time_au, const State& s,
                 bool vel, bool img) {
  const Units& U = c.units;
  float lf = (float)U.factor(LENGTH), vf = (float)U.factor(VELOCITY);
  std::string props = std::string("species:S:1:pos:R:3") + (vel ? ":velo:R:3" : "") + (img ? ":image:I:3" : "");
  float L[6];
  for (int k = 0; k < 6; ++k) L[k] = (float)s.lattice[k] * lf;
  std::string z = fmt_e(0.0);
  std::string buf = std::to_string(s.n) + "\n";
  buf += "Lattice=\"" + fmt_e(L[0]) + " " + z + " " + z + " " + fmt_e(L[3]) + " " + fmt_e(L[1]) + " " +
         z + " " + fmt_e(L[4]) + " " + fmt_e(L[5]) + " " + fmt_e(L[2]) + "\" Properties=" + props +
         " Step=" + std::to_string(step) + " Time=" + fmt_e(U.to_user(TIME, time_au)) + "\n";
  for (uint64_t i = 0; i < s.n; ++i) {
    buf += c.type_names[s.type[i]];
    for (int d = 0; d < 3; ++d) buf += " " + fmt_e(s.posq[4 * i + d] * lf);
    if (vel) {
      buf += " " + fmt_e(s.vx[i] * vf);
      buf += " " + fmt_e(s.vy[i] * vf);
      buf += " " + fmt_e(s.vz[i] * vf);
    }
    if (img)
      buf += " " + std::to_string(s.ix[i]) + " " + std::to_string(s.iy[i]) + " " + std::to_string(s.iz[i]);
    buf += "\n";
    if (buf.size() > (1u << 20)) {
      o.put(buf);
      buf.clear();
    }
  }
  o.put(buf);
}

struct StageClock {
  struct Acc { uint64_t count = 0; double total = 0, mn = 1e300, mx = 0; };
  std::vector<std::pair<std::string, Acc>> acc;
  void add(const std::string& name, double ms) {
    for (auto& a : acc)
      if (a.first == name) {
        a.second.count++; a.second.total += ms; a.second.mn = std::min(a.second.mn, ms);
        a.second.mx = std::max(a.second.mx, ms);
        return;
      }
    Acc a;
    a.count = 1; a.total = ms; a.mn = ms; a.mx = ms;
    acc.emplace_back(name, a);
  }
  void write(const std::string& path) const {
    OutFile o;
    o.open(path);
    char line[256];
    snprintf(line, sizeof line, "%-28s %10s %14s %13s %11s %11s\n", "stage", "count", "total_ms",
             "mean_us", "min_us", "max_us");
    o.put(line);
    for (auto& a : acc) {
      snprintf(line, sizeof line, "%-28s %10" PRIu64 " %14.3f %13.1f %11.1f %11.1f\n",
               a.first.c_str(), a.second.count, a.second.total,
               1e3 * a.second.total / (double)a.second.count, 1e3 * a.second.mn, 1e3 * a.second.mx);
      o.put(line);
    }
    o.close();
  }
};

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// ---- the engine for one phase ------------------------------------------------
struct Engine {
  hd_ctx* ctx = nullptr;
  ~Engine() {
    if (ctx) hd_ctx_destroy(ctx);
  }
  void check(int rc, const char* what, int code) {
    if (rc != HD_OK) {
      std::string m = std::string(what) + ": " + (ctx ? hd_last_error(ctx) : "no context");
      throw Failure{code, m};
    }
  }
};

double param(const Slot& s, const char* key, Dim d, const Units& U, const double* dflt = nullptr) {
  return U.to_au(d, get_num(s.params, key, s.kind, dflt));
}

// Builds and sets up a context for the phase: the particle state, the
// force field, the topology and the phase's slots.
void build_engine(Engine& e, const Config& c, const State& s, const Topology& t, const PhaseCfg& ph) {
  const Units& U = c.units;
  e.check(hd_ctx_create(0, &e.ctx), "GPU initialisation", 1);
  // exclusion keys: one per constraint group (its intramolecular pairs
  // are excluded by the key rule), a unique key per other atom
  std::vector<uint32_t> key(s.n);
  for (uint64_t i = 0; i < s.n; ++i) key[i] = (uint32_t)(t.groups.size() + i);
  for (size_t g = 0; g < t.groups.size(); ++g)
    for (uint32_t a : t.groups[g]) key[a] = (uint32_t)g;
  e.check(hd_set_system(e.ctx, s.n, s.posq.data(), s.mass.data(), s.type.data(), key.data()),
          "particle upload", 1);
  e.check(hd_set_lattice(e.ctx, s.lattice), "box", 1);
  e.check(hd_set_images(e.ctx, s.ix.data(), s.iy.data(), s.iz.data()), "images", 1);
  e.check(hd_set_velocities(e.ctx, s.vx.data(), s.vy.data(), s.vz.data()), "velocities", 1);
  int nt = c.ntypes();
  e.check(hd_set_lj(e.ctx, nt, c.sigma.data(), c.eps.data(), c.cutoff.data(), c.rswitch.data()),
          "pair interactions", 1);
  if (c.have_spme) {
    if (c.spline_order != 4 && c.spline_order != 5)
      fail1("spme.spline_order: orders 4 and 5 are built");
    e.check(hd_set_spme(e.ctx, c.alpha, c.r_cut_real, c.grid[0], c.grid[1], c.grid[2],
                        c.spline_order, 1.0),
            "spme", 1);
  }
  if (c.all_pairs) fail1("neighbor_list.mode = \"all-pairs\" is not built (use cell-list)");
  if (c.r_skin >= 0) e.check(hd_set_neighbor(e.ctx, c.r_skin), "neighbor_list", 1);
  // constraints: SETTLE groups by geometry (the type's d_OH, d_HH and the
  // group's O and H masses, settle.rs:196-240) and SHAKE groups
  std::vector<std::array<double, 4>> skeys;
  std::vector<std::vector<uint32_t>> settle;
  std::vector<uint32_t> goff{0}, gatoms, coff{0};
  std::vector<uint8_t> ci, cj;
  std::vector<double> cd;
  for (size_t g = 0; g < t.groups.size(); ++g) {
    const ConstraintType& ct = c.ctypes[t.group_type[g]];
    if (ct.kind == "settle") {
      const auto& at = t.groups[g];
      double m_o = s.mass[at[0]], m_h1 = s.mass[at[1]], m_h2 = s.mass[at[2]];
      if (m_h1 != m_h2)
        fail1("settle group " + std::to_string(g) + ": the two hydrogens must have equal masses");
      std::array<double, 4> key = {ct.d_oh, ct.d_hh, m_o, m_h1};
      size_t k = 0;
      while (k < skeys.size() && skeys[k] != key) ++k;
      if (k == skeys.size()) {
        skeys.push_back(key);
        settle.emplace_back();
      }
      for (uint32_t a : at) settle[k].push_back(a);
    } else {
      for (uint32_t a : t.groups[g]) gatoms.push_back(a);
      goff.push_back((uint32_t)gatoms.size());
      for (auto& k : ct.cons) {
        ci.push_back((uint8_t)std::get<0>(k));
        cj.push_back((uint8_t)std::get<1>(k));
        cd.push_back(std::get<2>(k));
      }
      coff.push_back((uint32_t)ci.size());
    }
  }
  for (size_t k = 0; k < skeys.size(); ++k) {
    const auto& q = skeys[k];
    auto setter = k == 0 ? hd_set_settle : hd_add_settle;
    e.check(setter(e.ctx, settle[k].size() / 3, settle[k].data(), q[0], q[1], q[2], q[3]), "settle",
            1);
  }
  if (goff.size() > 1)
    e.check(hd_set_shake(e.ctx, goff.size() - 1, goff.data(), gatoms.data(), coff.data(), ci.data(),
                         cj.data(), cd.data()),
            "shake", 1);
  if (!t.angle_k.empty())
    e.check(hd_set_angles(e.ctx, t.angle_k.size(), t.angle_ijk.data(), t.angle_k.data(), t.angle_t0.data()),
            "angles", 1);
  if (!t.bond_kind.empty())
    e.check(hd_set_bonds(e.ctx, t.bond_kind.size(), t.bond_ij.data(), t.bond_kind.data(), t.bond_p.data()),
            "bonds", 1);
  if (!t.dih_k.empty())
    e.check(hd_set_dihedrals(e.ctx, t.dih_k.size(), t.dih_ijkl.data(), t.dih_k.data(), t.dih_n.data(),
                             t.dih_phi0.data()),
            "dihedrals", 1);
  if (!t.excl.empty()) {
    std::vector<uint32_t> ij;
    std::vector<double> slj, sc;
    for (auto& kv : t.excl) {
      ij.push_back(kv.first.first);
      ij.push_back(kv.first.second);
      slj.push_back(kv.second.first);
      sc.push_back(kv.second.second);
    }
    e.check(hd_set_exclusions(e.ctx, slj.size(), ij.data(), slj.data(), sc.data()), "exclusions", 1);
  }
  // slots
  if (ph.minimize) {
    e.check(hd_set_integrator(e.ctx, 1.0, 0), "minimizer", 1);
  } else {
    const Slot& in = ph.integrator;
    if (in.kind == "velocity-verlet") {
      check_keys(in.params, {"kind", "lossless"}, "phase.integrator");
      e.check(hd_set_integrator(e.ctx, ph.dt, get_bool(in.params, "lossless", false) ? 1 : 0),
              "integrator", 1);
    } else if (in.kind == "langevin-baoab") {
      check_keys(in.params, {"kind", "friction", "temperature", "seed"}, "phase.integrator");
      e.check(hd_set_integrator(e.ctx, ph.dt, 0), "integrator", 1);
      e.check(hd_set_langevin(e.ctx, param(in, "friction", INV_TIME, U), param(in, "temperature", TEMPERATURE, U),
                              get_u64(in.params, "seed", "phase.integrator")),
              "integrator", 1);
    } else if (in.kind == "mtk-npt") {
      check_keys(in.params, {"kind", "temperature", "pressure", "tau_t", "tau_p", "chain_length",
                             "yoshida_order", "n_resp"},
                 "phase.integrator");
      uint64_t d3 = 3, d1 = 1;
      e.check(hd_set_integrator(e.ctx, ph.dt, 0), "integrator", 1);
      e.check(hd_set_mtk(e.ctx, param(in, "temperature", TEMPERATURE, U),
                         param(in, "pressure", PRESSURE, U), param(in, "tau_t", TIME, U),
                         param(in, "tau_p", TIME, U),
                         (int)get_u64(in.params, "chain_length", "phase.integrator", &d3),
                         (int)get_u64(in.params, "yoshida_order", "phase.integrator", &d3),
                         (int)get_u64(in.params, "n_resp", "phase.integrator", &d1), t.n_constraints),
              "integrator", 1);
    } else {
      fail1("phase.integrator: kind `" + in.kind +
            "` is not built (velocity-verlet, langevin-baoab, mtk-npt)");
    }
    const Slot& th = ph.thermostat;
    if (!th.kind.empty()) {
      if (th.kind == "nose-hoover-chain") {
        check_keys(th.params, {"kind", "temperature", "tau", "chain_length", "yoshida_order", "n_resp"},
                   "phase.thermostat");
        uint64_t d3 = 3, d1 = 1;
        e.check(hd_set_nhc(e.ctx, param(th, "temperature", TEMPERATURE, U), param(th, "tau", TIME, U),
                           (int)get_u64(th.params, "chain_length", "phase.thermostat", &d3),
                           (int)get_u64(th.params, "yoshida_order", "phase.thermostat", &d3),
                           (int)get_u64(th.params, "n_resp", "phase.thermostat", &d1), t.n_constraints),
                "thermostat", 1);
      } else if (th.kind == "csvr") {
        check_keys(th.params, {"kind", "temperature", "tau", "seed"}, "phase.thermostat");
        e.check(hd_set_csvr(e.ctx, param(th, "temperature", TEMPERATURE, U), param(th, "tau", TIME, U),
                            get_u64(th.params, "seed", "phase.thermostat"), t.n_constraints),
                "thermostat", 1);
      } else if (th.kind == "berendsen") {
        check_keys(th.params, {"kind", "temperature", "tau"}, "phase.thermostat");
        e.check(hd_set_berendsen(e.ctx, param(th, "temperature", TEMPERATURE, U), param(th, "tau", TIME, U),
                                 t.n_constraints),
                "thermostat", 1);
      } else if (th.kind == "andersen") {
        check_keys(th.params, {"kind", "temperature", "collision_rate", "seed"}, "phase.thermostat");
        e.check(hd_set_andersen(e.ctx, param(th, "temperature", TEMPERATURE, U),
                                param(th, "collision_rate", INV_TIME, U),
                                get_u64(th.params, "seed", "phase.thermostat")),
                "thermostat", 1);
      } else {
        fail1("phase.thermostat: kind `" + th.kind +
              "` is not built (nose-hoover-chain, csvr, berendsen, andersen)");
      }
    }
    const Slot& ba = ph.barostat;
    if (!ba.kind.empty()) {
      if (ba.kind == "c-rescale") {
        check_keys(ba.params, {"kind", "pressure", "temperature", "tau", "compressibility", "seed"},
                   "phase.barostat");
        e.check(hd_set_crescale(e.ctx, param(ba, "pressure", PRESSURE, U), param(ba, "tau", TIME, U),
                                param(ba, "compressibility", INV_PRESSURE, U),
                                param(ba, "temperature", TEMPERATURE, U),
                                get_u64(ba.params, "seed", "phase.barostat")),
                "barostat", 1);
      } else if (ba.kind == "berendsen") {
        check_keys(ba.params, {"kind", "pressure", "tau", "compressibility"}, "phase.barostat");
        e.check(hd_set_berendsen_barostat(e.ctx, param(ba, "pressure", PRESSURE, U), param(ba, "tau", TIME, U),
                                          param(ba, "compressibility", INV_PRESSURE, U)),
                "barostat", 1);
      } else if (ba.kind == "monte-carlo") {
        // host-orchestrated moves between steps (run_md); the device needs
        // only the molecule partition
        check_keys(ba.params, {"kind", "pressure", "temperature", "frequency", "volume_step", "seed"},
                   "phase.barostat");
        if (!t.mol_off.empty())
          e.check(hd_set_molecules(e.ctx, t.mol_off.size() - 1, t.mol_off.data(), t.mol_atoms.data()),
                  "molecules", 1);
      } else {
        fail1("phase.barostat: kind `" + ba.kind + "` is not built (c-rescale, berendsen, monte-carlo)");
      }
    }
  }
  e.check(hd_setup(e.ctx), "setup", 1);
}

void download(Engine& e, State& s, int code) {
  e.check(hd_get_positions(e.ctx, s.posq.data(), s.ix.data(), s.iy.data(), s.iz.data()), "download", code);
  e.check(hd_get_velocities(e.ctx, s.vx.data(), s.vy.data(), s.vz.data()), "download", code);
  e.check(hd_get_lattice(e.ctx, s.lattice), "download", code);
}

uint64_t thermal_dof(const State& s, const Topology& t) {
  long long d = 3LL * (long long)s.n - (long long)t.n_constraints - 3;
  return d > 0 ? (uint64_t)d : 0;
}

// log columns beyond step,time,kinetic_energy,temperature: the thermostat's
// conserved quantity, then the barostat's pressure, volume and conserved
// quantity (the log_column_names of each slot)
std::vector<std::pair<std::string, Dim>> extra_columns(const PhaseCfg& ph) {
  std::vector<std::pair<std::string, Dim>> cols;
  if (ph.thermostat.kind == "nose-hoover-chain") cols.push_back({"nhc_conserved", ENERGY});
  if (ph.thermostat.kind == "csvr") cols.push_back({"csvr_conserved", ENERGY});
  if (ph.thermostat.kind == "berendsen") cols.push_back({"berendsen_conserved", ENERGY});
  if (ph.thermostat.kind == "andersen") cols.push_back({"andersen_conserved", ENERGY});
  if (ph.integrator.kind == "mtk-npt") {
    cols.push_back({"pressure", PRESSURE});
    cols.push_back({"box_volume", NONE});
    cols.push_back({"mtk_npt_conserved", ENERGY});
  }
  if (ph.barostat.kind == "c-rescale" || ph.barostat.kind == "berendsen") {
    cols.push_back({"pressure", PRESSURE});
    cols.push_back({"box_volume", NONE});
    if (ph.barostat.kind == "c-rescale") cols.push_back({"c_rescale_conserved", ENERGY});
  } else if (ph.barostat.kind == "monte-carlo") {
    cols.push_back({"box_volume", NONE});
    cols.push_back({"mc_acceptance", NONE});
    cols.push_back({"mc_volume_step", NONE});
    cols.push_back({"mc_conserved", ENERGY});
  }
  return cols;
}

// Monte-Carlo barostat bookkeeping on the host (mc_barostat.rs:122-150,
// 353-385): draw counter, acceptance, the adaptive step every 10 attempts
struct McState {
  bool on = false;
  double pressure = 0, kt = 0, max_dv = 0, injection = 0, volume = 0;
  uint64_t freq = 25, seed = 0, counter = 0, attempted = 0, accepted = 0, n_att = 0, n_acc = 0;
  double acceptance() const { return attempted ? (double)accepted / (double)attempted : 0.0; }
};

void log_row(OutFile& o, const Config& c, const PhaseCfg& ph, Engine& e, uint64_t step, uint64_t ndof,
             const std::vector<std::pair<std::string, Dim>>& cols, const McState& mc) {
  double sc[8];
  e.check(hd_get_scalars(e.ctx, sc), "scalars", 2);
  double ke = sc[0], pe = sc[1];
  double T = ndof ? 2.0 * ke / (double)ndof : 0.0;
  const Units& U = c.units;
  std::string row = std::to_string(step) + "," + fmt_e(U.to_user(TIME, (double)step * ph.dt)) + "," +
                    fmt_e(U.to_user(ENERGY, ke)) + "," + fmt_e(U.to_user(TEMPERATURE, T));
  double baro[5] = {0, 0, 0, 0, 0};
  if (!ph.barostat.kind.empty() && !mc.on) e.check(hd_get_barostat(e.ctx, baro), "barostat", 2);
  double mtk[5] = {0, 0, 0, 0, 0};
  if (ph.integrator.kind == "mtk-npt") {
    e.check(hd_get_mtk(e.ctx, mtk), "integrator", 2);
    baro[1] = mtk[0];
    baro[2] = mtk[1];
  }
  double p0 = 0.0;
  if (ph.barostat.kind == "c-rescale") p0 = param(ph.barostat, "pressure", PRESSURE, U);
  for (auto& col : cols) {
    double v = 0;
    if (col.first == "pressure") v = baro[1];
    else if (col.first == "box_volume") v = mc.on ? mc.volume : baro[2];
    else if (col.first == "c_rescale_conserved") v = ke + pe + p0 * baro[2] - baro[3];
    else if (col.first == "mc_acceptance") v = mc.acceptance();
    else if (col.first == "mc_volume_step") v = mc.max_dv;
    else if (col.first == "mc_conserved") v = ke + pe + mc.pressure * mc.volume - mc.injection;
    else if (col.first == "mtk_npt_conserved") v = ke + pe + mtk[2];
    else v = ke + pe + sc[5];  // thermostat conserved quantity
    row += "," + fmt_e(U.to_user(col.second, v));
  }
  o.put(row + "\n");
}

void run_md(const Config& c, State& s, const Topology& t, const PhaseCfg& ph, uint64_t& frames,
            uint64_t& rows) {
  StageClock clock;
  double t0 = now_ms();
  Engine e;
  build_engine(e, c, s, t, ph);
  clock.add("setup", now_ms() - t0);
  if (!s.have_velocities) {
    // first MD phase from an init file without velocities: they were
    // sampled on the host; project them onto the constraints once
    e.check(hd_constrain_velocities(e.ctx), "velocity projection", 1);
    s.have_velocities = true;
  }
  OutFile traj, log;
  if (ph.traj_every) traj.open(ph.traj_path);
  if (ph.log_every) log.open(ph.log_path);
  uint64_t ndof = thermal_dof(s, t);
  auto cols = extra_columns(ph);
  McState mc;
  if (ph.barostat.kind == "monte-carlo") {
    const Units& U = c.units;
    uint64_t d25 = 25;
    mc.on = true;
    mc.pressure = param(ph.barostat, "pressure", PRESSURE, U);
    mc.kt = param(ph.barostat, "temperature", TEMPERATURE, U);
    mc.freq = get_u64(ph.barostat.params, "frequency", "phase.barostat", &d25);
    if (mc.freq == 0) fail1("phase.barostat.frequency must be >= 1");
    mc.seed = get_u64(ph.barostat.params, "seed", "phase.barostat");
    mc.volume = s.lattice[0] * s.lattice[1] * s.lattice[2];
    double dflt = 0.01 * mc.volume;
    mc.max_dv = get_num(ph.barostat.params, "volume_step", "phase.barostat", &dflt);
  }
  if (ph.log_every) {
    std::string head = "step,time,kinetic_energy,temperature";
    for (auto& col : cols) head += "," + col.first;
    log.put(head + "\n");
  }
  // forces, energies and the kinetic energy of the start
  e.check(hd_force_eval(e.ctx, 1), "force evaluation", 2);
  if (ph.log_every) {
    log_row(log, c, ph, e, 0, ndof, cols, mc);
    ++rows;
  }
  if (ph.traj_every) {
    download(e, s, 2);
    write_frame(traj, c, 0, 0.0, s, ph.include_velocities, ph.include_images);
    ++frames;
  }
  uint64_t step = 0;
  while (step < ph.n_steps) {
    uint64_t next = ph.n_steps;
    if (ph.log_every) next = std::min(next, (step / ph.log_every + 1) * ph.log_every);
    if (ph.traj_every) next = std::min(next, (step / ph.traj_every + 1) * ph.traj_every);
    if (mc.on) next = std::min(next, (step / mc.freq + 1) * mc.freq);
    bool want_log = ph.log_every && next % ph.log_every == 0;
    bool want_frame = ph.traj_every && next % ph.traj_every == 0;
    double ts = now_ms();
    e.check(hd_step(e.ctx, (int)(next - step), want_log ? 1 : 0), "step", 2);
    e.check(hd_sync(e.ctx), "step", 2);
    clock.add("md_steps", now_ms() - ts);
    step = next;
    if (mc.on && step % mc.freq == 0) {
      // the move follows the dynamics step and precedes its output
      // (runner.rs, rq-03a5a290)
      ts = now_ms();
      double out[4];
      mc.counter += 1;
      e.check(hd_mc_barostat_move(e.ctx, mc.pressure, mc.kt, mc.max_dv, mc.seed, mc.counter, out),
              "monte-carlo barostat", 2);
      bool acc = out[0] != 0.0;
      mc.attempted++;
      mc.n_att++;
      if (acc) {
        mc.accepted++;
        mc.n_acc++;
        mc.injection += mc.pressure * (out[2] - out[1]);
      }
      mc.volume = out[2];
      if (mc.n_att >= 10) {
        if ((double)mc.n_acc < 0.25 * (double)mc.n_att) mc.max_dv /= 1.1;
        else if ((double)mc.n_acc > 0.75 * (double)mc.n_att)
          mc.max_dv = std::min(mc.max_dv * 1.1, 0.3 * mc.volume);
        mc.n_att = mc.n_acc = 0;
      }
      clock.add("mc_barostat", now_ms() - ts);
    }
    if (want_log) {
      ts = now_ms();
      log_row(log, c, ph, e, step, ndof, cols, mc);
      ++rows;
      clock.add("log_write", now_ms() - ts);
    }
    if (want_frame) {
      ts = now_ms();
      download(e, s, 2);
      write_frame(traj, c, step, (double)step * ph.dt, s, ph.include_velocities, ph.include_images);
      ++frames;
      clock.add("trajectory_write", now_ms() - ts);
    }
  }
  download(e, s, 2);
  traj.close();
  log.close();
  clock.write(ph.timings_path);
}

const char* run_minimize(const Config& c, State& s, const Topology& t, const PhaseCfg& ph,
                         uint64_t& frames, uint64_t& rows, uint64_t& iters) {
  StageClock clock;
  double t0 = now_ms();
  Engine e;
  build_engine(e, c, s, t, ph);
  clock.add("setup", now_ms() - t0);
  OutFile traj, minlog;
  if (ph.traj_every) traj.open(ph.traj_path);
  if (ph.minlog_every) {
    minlog.open(ph.minlog_path);
    minlog.put("iter,energy,max_force,step,accepted\n");
  }
  const Units& U = c.units;
  double st[4];  // energy, max force, step used, accepted
  e.check(hd_minimize_begin(e.ctx, ph.sd_initial, st), "minimization", 2);
  double energy = st[0], fmax = st[1];
  auto row = [&](uint64_t it, double en, double fm, double stp, bool acc) {
    minlog.put(std::to_string(it) + "," + fmt_e(U.to_user(ENERGY, en)) + "," + fmt_e(U.to_user(FORCE, fm)) +
               "," + fmt_e(U.to_user(LENGTH, stp)) + "," + (acc ? "1" : "0") + "\n");
    ++rows;
  };
  if (ph.minlog_every) row(0, energy, fmax, ph.sd_initial, true);
  if (ph.traj_every) {
    download(e, s, 2);
    write_frame(traj, c, 0, 0.0, s, false, ph.include_images);
    ++frames;
  }
  const char* reason = nullptr;
  for (uint64_t it = 1; it <= ph.sd_max_iter && !reason; ++it) {
    double ts = now_ms();
    double prev = energy;
    double out[4];
    e.check(hd_minimize_step(e.ctx, ph.sd_max, ph.sd_inc, ph.sd_dec, out), "minimization", 2);
    clock.add("sd_iteration", now_ms() - ts);
    bool acc = out[3] != 0.0;
    energy = out[0];
    fmax = out[1];
    iters = it;
    if (ph.minlog_every && it % ph.minlog_every == 0) row(it, energy, fmax, out[2], acc);
    if (ph.traj_every && it % ph.traj_every == 0) {
      download(e, s, 2);
      write_frame(traj, c, it, 0.0, s, false, ph.include_images);
      ++frames;
    }
    // convergence (src/minimizer/steepest_descent.rs check_convergence)
    if (fmax == 0.0) reason = "force_zero";
    else if (ph.sd_ftol > 0 && fmax <= ph.sd_ftol) reason = "force_tolerance";
    else if (acc && ph.sd_etol > 0 &&
             std::fabs(energy - prev) <= ph.sd_etol * std::max(std::max(std::fabs(energy), std::fabs(prev)), 1e-30))
      reason = "energy_tolerance";
  }
  if (!reason)
    fail2("minimization `" + ph.name + "` reached max_iterations (" + std::to_string(ph.sd_max_iter) +
          ") without converging");
  download(e, s, 2);
  traj.close();
  minlog.close();
  clock.write(ph.timings_path);
  return reason;
}

std::string fmt_elapsed(double ms) {
  char b[64];
  if (ms < 10.0) snprintf(b, sizeof b, "%.0f \xC2\xB5s", ms * 1e3);
  else snprintf(b, sizeof b, "%.0f ms", ms);
  return b;
}

std::vector<std::string> output_paths(const Config& c) {
  std::vector<std::string> out;
  for (auto& p : c.phases) {
    if (p.minimize) {
      if (p.minlog_every) out.push_back(p.minlog_path);
    } else if (p.log_every) {
      out.push_back(p.log_path);
    }
    if (p.traj_every) out.push_back(p.traj_path);
    out.push_back(p.timings_path);
  }
  return out;
}

void check_box(const Config& c, const State& s, const Topology& t) {
  (void)t;
  double need = 3.0 * (c.max_cutoff() + (c.r_skin >= 0 ? c.r_skin : 0.3 * c.max_cutoff()));
  // widths perpendicular to the other two box vectors (min_perpendicular_width)
  const double* L = s.lattice;
  double ilx = 1.0 / L[0], ily = 1.0 / L[1], ilz = 1.0 / L[2];
  double a1 = -L[3] * ilx * ily, a2 = (L[3] * L[5] - L[4] * L[1]) * ilx * ily * ilz;
  double b2 = -L[5] * ily * ilz;
  double w[3] = {1.0 / sqrt(ilx * ilx + a1 * a1 + a2 * a2), 1.0 / sqrt(ily * ily + b2 * b2), L[2]};
  for (int d = 0; d < 3; ++d)
    if (w[d] < need) {
      char b[200];
      snprintf(b, sizeof b,
               "simulation box perpendicular width along lattice direction `%c` is %g, below the required %g",
               "abc"[d], c.units.to_user(LENGTH, w[d]), c.units.to_user(LENGTH, need));
      fail1(b);
    }
}

int cmd_run(const std::string& path) {
  Config c = load_config(path);
  for (auto& p : output_paths(c))
    if (file_exists(p)) fail1("output file already exists: `" + p + "`");
  State s = load_init(c);
  Topology t = load_topology(c, s);
  check_box(c, s, t);
  if (!s.have_velocities) maxwell_boltzmann(s, s.mass_d, c.temperature, c.seed, t.n_constraints);
  double total = 0;
  uint64_t total_n = 0;
  for (auto& ph : c.phases) {
    uint64_t frames = 0, rows = 0;
    double t0 = now_ms();
    if (ph.minimize) {
      uint64_t iters = 0;
      const char* why = run_minimize(c, s, t, ph, frames, rows, iters);
      double ms = now_ms() - t0;
      total += ms;
      total_n += iters;
      printf("[heddlemd] phase `%s`: %" PRIu64 " iters in %s (converged: %s, frames: %" PRIu64
             ", log rows: %" PRIu64 ")\n",
             ph.name.c_str(), iters, fmt_elapsed(ms).c_str(), why, frames, rows);
    } else {
      run_md(c, s, t, ph, frames, rows);
      double ms = now_ms() - t0;
      total += ms;
      total_n += ph.n_steps;
      printf("[heddlemd] phase `%s`: %" PRIu64 " steps in %s (frames: %" PRIu64 ", log rows: %" PRIu64 ")\n",
             ph.name.c_str(), ph.n_steps, fmt_elapsed(ms).c_str(), frames, rows);
    }
    fflush(stdout);
  }
  printf("[heddlemd] complete: %zu phases, %" PRIu64 " steps in %s\n", c.phases.size(), total_n,
         fmt_elapsed(total).c_str());
  return 0;
}

int cmd_lint(const std::string& path, bool with_gpu) {
  std::vector<std::pair<std::string, std::string>> report = {
      {"config", "skipped (earlier check failed)"}, {"output paths", "skipped (earlier check failed)"},
      {"init", "skipped (earlier check failed)"}, {"box/cutoff", "skipped (earlier check failed)"},
      {"topology", "skipped (earlier check failed)"},
      {"gpu", with_gpu ? "skipped (earlier check failed)" : "not checked (re-run with --with-gpu)"}};
  auto print = [&](bool ok) {
    printf("[heddlemd lint] %s\n", ok ? "OK" : "FAIL");
    for (auto& r : report) printf("  %-12s %s\n", r.first.c_str(), r.second.c_str());
    fflush(stdout);
  };
  size_t stage = 0;
  try {
    Config c = load_config(path);
    report[0].second = path;
    stage = 1;
    for (auto& p : output_paths(c))
      if (file_exists(p)) fail1("output file already exists: `" + p + "`");
    report[1].second = "none pre-exist";
    stage = 2;
    State s = load_init(c);
    char b[256];
    snprintf(b, sizeof b, "resolved, %" PRIu64 " particles, box %s x %s x %s", s.n,
             fmt_e(c.units.to_user(LENGTH, s.lattice[0]), 1).c_str(),
             fmt_e(c.units.to_user(LENGTH, s.lattice[1]), 1).c_str(),
             fmt_e(c.units.to_user(LENGTH, s.lattice[2]), 1).c_str());
    report[2].second = b;
    stage = 3;
    Topology t;
    check_box(c, s, t);
    report[3].second = "ok";
    stage = 4;
    t = load_topology(c, s);
    if (c.topology_path.empty()) {
      report[4].second = "not supplied";
    } else {
      snprintf(b, sizeof b, "%zu bonds, %zu angles, %zu dihedrals, %zu constraint groups",
               t.bond_kind.size(), t.angle_k.size(), t.dih_k.size(), t.groups.size());
      report[4].second = b;
    }
    stage = 5;
    if (with_gpu) {
      if (!s.have_velocities) maxwell_boltzmann(s, s.mass_d, c.temperature, c.seed, t.n_constraints);
      for (auto& ph : c.phases) {
        Engine e;
        build_engine(e, c, s, t, ph);
      }
      report[5].second = "ok";
    }
  } catch (const Failure& f) {
    report[stage].second = "FAIL";
    print(false);
    fprintf(stderr, "error: %s\n", f.msg.c_str());
    return 1;
  }
  print(true);
  return 0;
}

void usage() {
  fprintf(stderr,
          "usage: heddlemd run     <config-path>\n"
          "       heddlemd lint    <config-path> [--with-gpu]\n"
          "       heddlemd analyze <analysis-path>\n");
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    usage();
    return 1;
  }
  std::string cmd = argv[1], path = argv[2];
  try {
    if (cmd == "run" && argc == 3) return cmd_run(path);
    if (cmd == "lint" && (argc == 3 || (argc == 4 && std::string(argv[3]) == "--with-gpu")))
      return cmd_lint(path, argc == 4);
    if (cmd == "analyze" && argc == 3) fail1("`analyze` is not built in this front end");
  } catch (const Failure& f) {
    fprintf(stderr, "error: %s\n", f.msg.c_str());
    return f.code;
  } catch (const std::exception& ex) {
    fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  usage();
  return 1;
}
