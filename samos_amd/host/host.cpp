// Host-side plugin classes, registries, configuration grammar, input reader
// and observers.  Parameter names, defaults and error texts follow the
// reference constructors cited at each class; computation is forwarded to the
// device through include/samos_hip.h.
#include "host.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <iostream>
#include <sstream>

namespace samos_host {

// ---------------------------------------------------------------------------
Messenger::Messenger(const std::string& file) {
  if (!file.empty()) set_file(file);
}

void Messenger::set_file(const std::string& file) {
  out_.open(file);
  to_file_ = out_.good();
}

void Messenger::msg(Level l, const std::string& text) {
  static const char* tag[] = {"INFO", "WARNING", "ERROR"};
  std::time_t now = std::time(nullptr);
  char buf[64];
  std::strftime(buf, sizeof buf, "%Y-%m-%d %H:%M:%S", std::localtime(&now));
  std::ostream& o = to_file_ ? static_cast<std::ostream&>(out_) : (l == ERROR ? std::cerr : std::cout);
  o << buf << " : " << tag[l] << " : " << text << std::endl;
  if (to_file_ && l == ERROR) std::cerr << tag[l] << " : " << text << std::endl;
}

void fail(const MessengerPtr& msg, const std::string& error_line, const std::string& what) {
  if (msg) msg->msg(Messenger::ERROR, error_line);
  throw std::runtime_error(what);
}

// ---------------------------------------------------------------------------
// Grammar.  A command line is `command rest`; the parameter block is
// `{ key [= value] ((; | ,) key [= value])* [;]* }` with keys
// [a-zA-Z_][a-zA-Z_0-9]* and values [-a-zA-Z_0-9+./]+
// (src/parser/parse_parameters.hpp:45-59).
// ---------------------------------------------------------------------------
static std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

bool parse_command(const std::string& line, CommandLine& out) {
  std::string t = trim(line);
  if (t.empty() || t[0] == '#') return false;
  size_t i = 0;
  while (i < t.size() && (std::isalnum((unsigned char)t[i]) || t[i] == '_')) ++i;
  out.command = t.substr(0, i);
  out.rest = trim(t.substr(i));
  return !out.command.empty();
}

static bool is_key_start(char c) { return std::isalpha((unsigned char)c) || c == '_'; }
static bool is_key_char(char c) { return std::isalnum((unsigned char)c) || c == '_'; }
static bool is_value_char(char c) {
  return std::isalnum((unsigned char)c) || c == '-' || c == '_' || c == '+' || c == '.' || c == '/';
}

bool parse_params(const std::string& text, pairs_type& params) {
  std::string s = trim(text);
  size_t i = 0;
  auto skip = [&]() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; };
  skip();
  if (i >= s.size() || s[i] != '{') return false;
  ++i;
  skip();
  if (i < s.size() && s[i] == '}') { ++i; skip(); return i == s.size(); }
  bool first = true;
  while (true) {
    skip();
    if (!first) {
      if (i < s.size() && (s[i] == ';' || s[i] == ',')) {
        ++i;
        skip();
        while (i < s.size() && s[i] == ';') { ++i; skip(); }
        if (i < s.size() && s[i] == '}') break;
      } else {
        break;
      }
    }
    first = false;
    if (i >= s.size() || !is_key_start(s[i])) return false;
    size_t k0 = i;
    while (i < s.size() && is_key_char(s[i])) ++i;
    std::string key = s.substr(k0, i - k0);
    skip();
    std::string value;
    if (i < s.size() && s[i] == '=') {
      ++i;
      skip();
      size_t v0 = i;
      while (i < s.size() && is_value_char(s[i])) ++i;
      if (i == v0) return false;
      value = s.substr(v0, i - v0);
    }
    params[key] = value;
  }
  skip();
  if (i >= s.size() || s[i] != '}') return false;
  ++i;
  skip();
  return i == s.size();
}

bool parse_attrib_params(const std::string& rest, std::string& attrib, pairs_type& params) {
  std::string s = trim(rest);
  size_t b = s.find('{');
  if (b == std::string::npos) {
    attrib = s;
    return true;
  }
  attrib = trim(s.substr(0, b));
  return parse_params(s.substr(b), params);
}

double to_double(const std::string& s, const std::string& what) {
  try {
    size_t pos = 0;
    double v = std::stod(s, &pos);
    if (pos != s.size()) throw std::invalid_argument(s);
    return v;
  } catch (const std::exception&) {
    throw std::runtime_error("bad numeric value \"" + s + "\" for " + what);
  }
}

int to_int(const std::string& s, const std::string& what) {
  try {
    size_t pos = 0;
    long v = std::stol(s, &pos);
    if (pos != s.size()) throw std::invalid_argument(s);
    return (int)v;
  } catch (const std::exception&) {
    throw std::runtime_error("bad integer value \"" + s + "\" for " + what);
  }
}

static bool has(const pairs_type& p, const char* k) { return p.find(k) != p.end(); }

// ---------------------------------------------------------------------------
// System (src/system/system.cpp:64-452)
// ---------------------------------------------------------------------------
System::System(const std::string& input, MessengerPtr msg, const Box& box, bool periodic)
    : msg_(msg), box_(box), periodic_(periodic) {
  std::ifstream in(input);
  if (!in) fail(msg_, "Problem opening file " + input, "Problem opening file " + input);
  static const char* old_keys[] = {"id", "type", "radius", "x", "y", "z", "vx", "vy", "vz", "nx", "ny", "nz",
                                   "omega", "length", "ix", "iy", "iz"};
  std::map<std::string, int> col;
  bool has_keys = false;
  std::string line;
  std::vector<std::vector<std::string>> rows;
  while (std::getline(in, line)) {
    line = trim(line);
    std::transform(line.begin(), line.end(), line.begin(), [](unsigned char c) { return std::tolower(c); });
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    std::vector<std::string> tok;
    std::string w;
    while (ss >> w) tok.push_back(w);
    if (tok.empty()) continue;
    if (tok[0] == "keys:") {
      has_keys = true;
      for (size_t c = 1; c < tok.size(); ++c) col[tok[c]] = (int)c - 1;
      continue;
    }
    rows.push_back(tok);
  }
  if (!has_keys)
    for (int c = 0; c < 17; ++c) col[old_keys[c]] = c;
  n_ = (int)rows.size();
  std::map<std::string, std::vector<double>> d;
  const char* dnames[] = {"x", "y", "z", "vx", "vy", "vz", "nx", "ny", "nz", "omega", "radius", "mass"};
  for (const char* k : dnames) d[k].assign(n_, 0.0);
  d["radius"].assign(n_, 1.0);
  d["mass"].assign(n_, 1.0);
  std::vector<int> type(n_, 1), ix(n_, 0), iy(n_, 0), iz(n_, 0);
  for (int i = 0; i < n_; ++i) {
    const auto& r = rows[i];
    auto get = [&](const std::string& k, double& v) {
      auto it = col.find(k);
      if (it != col.end() && it->second < (int)r.size()) v = to_double(r[it->second], k);
    };
    for (const char* k : dnames) get(k, d[k][i]);
    double tv = 1.0;
    get("type", tv);
    type[i] = (int)tv;
    double fv = 0;
    fv = 0; get("ix", fv); ix[i] = (int)fv;
    fv = 0; get("iy", fv); iy[i] = (int)fv;
    fv = 0; get("iz", fv); iz[i] = (int)fv;
    const std::string id = std::to_string(i);
    if (d["x"][i] < box_.xlo() || d["x"][i] > box_.xhi())
      fail(msg_, "X coordinate of particle " + id + " is outside simulation box. Please update box size.",
           "Particle outside the box.");
    if (d["y"][i] < box_.ylo() || d["y"][i] > box_.yhi())
      fail(msg_, "Y coordinate of particle " + id + " is outside simulation box. Please update box size.",
           "Particle outside the box.");
    if (d["z"][i] < box_.zlo() || d["z"][i] > box_.zhi())
      fail(msg_, "Z coordinate of particle " + id + " is outside simulation box. Please update box size.",
           "Particle outside the box.");
    ntypes_ = std::max(ntypes_, type[i]);
  }
  int rc = samos_create(&ctx_, 0, n_, box_.Lx, box_.Ly, box_.Lz, periodic_ ? 1 : 0);
  if (rc != 0) fail(msg_, std::string("Device context: ") + samos_last_error(nullptr), "device context failed");
  const int fcode[] = {SAMOS_X, SAMOS_Y, SAMOS_Z, SAMOS_VX, SAMOS_VY, SAMOS_VZ, SAMOS_NX, SAMOS_NY, SAMOS_NZ,
                       SAMOS_OMEGA, SAMOS_RADIUS, SAMOS_MASS};
  for (int k = 0; k < 12; ++k) check(samos_set_field(ctx_, fcode[k], d[dnames[k]].data()));
  check(samos_set_field(ctx_, SAMOS_TYPE, type.data()));
  check(samos_set_field(ctx_, SAMOS_IX, ix.data()));
  check(samos_set_field(ctx_, SAMOS_IY, iy.data()));
  check(samos_set_field(ctx_, SAMOS_IZ, iz.data()));
  msg_->msg(Messenger::INFO, "Read data for " + std::to_string(n_) + " particles.");
  std::vector<int>& all = groups_["all"];
  for (int i = 0; i < n_; ++i) all.push_back(i);
}

void System::make_group(const std::string& name, pairs_type& p) {
  if (name == "boundary" || name == "tissue")
    fail(msg_, "Groups \"tissue\" and \"boundary\" are reserved and cannot be used.",
         "Trying to redefine reserved group name.");
  std::vector<char> add(n_, 1);
  auto crit = [&](const char* key, int field, bool is_int, int op) {
    // op: 0 equal, 1 not equal, 2 <=, 3 >=
    if (!p.count(key)) return;
    msg_->msg(Messenger::INFO, "Adding particle to group " + name + ".");
    if (is_int) {
      const int v = to_int(p[key], key);
      const std::vector<int> c = icolumn(field);
      for (int i = 0; i < n_; ++i) add[i] &= (op == 0 ? c[i] == v : c[i] != v) ? 1 : 0;
    } else {
      const double v = to_double(p[key], key);
      const std::vector<double> c = column(field);
      for (int i = 0; i < n_; ++i) add[i] &= (op == 0 ? c[i] == v : op == 2 ? c[i] <= v : c[i] >= v) ? 1 : 0;
    }
  };
  crit("type", SAMOS_TYPE, true, 0);
  crit("not_type", SAMOS_TYPE, true, 1);
  crit("r_eq", SAMOS_RADIUS, false, 0);
  crit("r_le", SAMOS_RADIUS, false, 2);
  crit("r_ge", SAMOS_RADIUS, false, 3);
  crit("x_le", SAMOS_X, false, 2);
  crit("x_ge", SAMOS_X, false, 3);
  crit("y_le", SAMOS_Y, false, 2);
  crit("y_ge", SAMOS_Y, false, 3);
  crit("z_le", SAMOS_Z, false, 2);
  crit("z_ge", SAMOS_Z, false, 3);
  std::vector<int> member(n_);
  std::vector<int>& g = groups_[name];
  std::vector<char> in(n_, 0);
  for (int i : g) in[i] = 1;
  for (int i = 0; i < n_; ++i) {
    member[i] = add[i];
    if (add[i] && !in[i]) g.push_back(i);
  }
  check(samos_group_define(ctx_, name.c_str(), member.data()));
  msg_->msg(Messenger::INFO, "Added " + std::to_string(g.size()) + " particles to group " + name + ".");
}

System::~System() {
  if (ctx_) samos_destroy(ctx_);
}

void System::check(int rc) const {
  if (rc != 0) {
    std::string e = samos_last_error(ctx_);
    fail(msg_, e, e);
  }
}

void System::set_integrator_step(double dt) {
  dt_ = dt;
  check(samos_set_timestep(ctx_, dt));
}

std::vector<double> System::column(int field) const {
  std::vector<double> v(n_);
  check(samos_get_field(ctx_, field, v.data()));
  return v;
}

std::vector<int> System::icolumn(int field) const {
  std::vector<int> v(n_);
  check(samos_get_field(ctx_, field, v.data()));
  return v;
}

// ---------------------------------------------------------------------------
// Value (src/system/value.hpp:53-170)
// ---------------------------------------------------------------------------
ValuePtr make_value(const std::string& kind, MessengerPtr msg, pairs_type& p) {
  auto v = std::make_shared<Value>();
  if (has(p, "min_val")) v->min_val = to_double(p["min_val"], "min_val");
  if (has(p, "max_val")) v->max_val = to_double(p["max_val"], "max_val");
  if (has(p, "steps")) v->steps = to_int(p["steps"], "steps");
  if (kind == "constant") {
    v->kind = SAMOS_VALUE_CONSTANT;
  } else if (kind == "linear") {
    v->kind = SAMOS_VALUE_LINEAR;
    if (v->max_val < v->min_val)
      fail(msg, "Value object. Minimum value has to be smaller than the maximum value.",
           "Value parameters are not properly set.");
    if (v->steps <= 0)
      fail(msg, "Value object. Number of steps has to be positive non-zero integer.",
           "Value parameters. Zero steps given.");
  } else {
    fail(msg, "Unknown value type " + kind + ".", "Unknown value type.");
  }
  return v;
}

// ---------------------------------------------------------------------------
// NeighbourList (src/system/neighbour_list.hpp:124-217)
// ---------------------------------------------------------------------------
NeighbourList::NeighbourList(SystemPtr sys, MessengerPtr msg, double cutoff, double pad, pairs_type& params)
    : sys_(sys), cut_(cutoff), pad_(pad) {
  for (const char* k : {"triangulation", "disable_nlist"})
    if (has(params, k))
      fail(msg, std::string("Neighbour list option ") + k + " (tissue models) is not supported by this implementation.",
           "Unsupported neighbour list option.");
  sys_->check(samos_nlist_create(sys_->ctx(), cutoff, pad));
  long long info[8];
  sys_->check(samos_nlist_info(sys_->ctx(), info));
  if (info[0])
    msg->msg(Messenger::INFO, "Using cell lists for neighbour list builds.");
  else
    msg->msg(Messenger::INFO, "Box dimensions are too small to be able to use cell lists. Neighbour list will be "
                              "built using N^2 algorithm.");
}

void NeighbourList::build() { sys_->check(samos_nlist_build(sys_->ctx())); }

// ---------------------------------------------------------------------------
// PairPotential (src/potentials/pair/pair_potential.hpp:62-129)
// ---------------------------------------------------------------------------
PairPotential::PairPotential(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, ValuePtr val, pairs_type& p)
    : sys_(sys), msg_(msg), nlist_(nlist), val_(val) {
  known_ = {"min_val", "max_val", "ntypes", "steps"};
  ntypes_ = sys_->ntypes();
  if (has(p, "ntypes")) ntypes_ = std::max(ntypes_, to_int(p["ntypes"], "ntypes"));
}

std::string PairPotential::params_ok(pairs_type& params) const {
  for (auto& kv : params)
    if (std::find(known_.begin(), known_.end(), kv.first) == known_.end()) return kv.first;
  return "";
}

static void need_types(const MessengerPtr& msg, pairs_type& pp, const std::string& where) {
  if (!has(pp, "type_1"))
    fail(msg, "type_1 has not been defined for pair potential parameters in " + where + ".",
         "Missing key for pair potential parameters.");
  if (!has(pp, "type_2"))
    fail(msg, "type_2 has not been defined for pair potential parameters in " + where + ".",
         "Missing key for pair potential parameters.");
}

// PairSoftPotential (src/potentials/pair/pair_soft_potential.hpp:63-180)
PairSoftPotential::PairSoftPotential(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, ValuePtr val,
                                     pairs_type& p)
    : PairPotential(sys, msg, nlist, val, p) {
  for (const char* k : {"k", "a", "use_particle_radii", "phase_in"}) known_.push_back(k);
  std::string bad = params_ok(p);
  if (!bad.empty())
    fail(msg_, "Parameter \"" + bad + "\" is not a valid parameter for soft pair potential.",
         "Unknown parameter \"" + bad + "\" in soft potential.");
  if (has(p, "phase_in"))
    fail(msg_, "Soft pair potential: phase_in (population growth) is not supported by this implementation.",
         "Unsupported parameter phase_in.");
  if (!has(p, "k")) msg_->msg(Messenger::WARNING, "No potential strength (k) specified for soft pair potential. Setting it to 1.");
  else k_ = to_double(p["k"], "k");
  if (!has(p, "a")) msg_->msg(Messenger::WARNING, "No potential range (a) specified for soft pair potential. Setting it to 2.");
  else a_ = to_double(p["a"], "a");
  const bool radii = has(p, "use_particle_radii");
  if (radii) msg_->msg(Messenger::WARNING, "Soft pair potential is set to use particle radii to control its range. Parameter a will be ignored.");
  sys_->check(samos_pair_soft(sys_->ctx(), k_, a_, radii ? 1 : 0));
}

void PairSoftPotential::set_pair_parameters(pairs_type& pp) {
  need_types(msg_, pp, "soft potential");
  const int t1 = to_int(pp["type_1"], "type_1"), t2 = to_int(pp["type_2"], "type_2");
  const double k = has(pp, "k") ? to_double(pp["k"], "k") : k_;
  const double a = has(pp, "a") ? to_double(pp["a"], "a") : a_;
  sys_->check(samos_pair_type_param(sys_->ctx(), SAMOS_PAIR_SOFT, t1, t2, k, a, 0.0, 0));
}

void PairSoftPotential::compute(double) { sys_->check(samos_pair_compute(sys_->ctx(), SAMOS_PAIR_SOFT, &energy_)); }

// PairLJPotential (src/potentials/pair/pair_lj_potential.hpp:61-248)
PairLJPotential::PairLJPotential(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, ValuePtr val, pairs_type& p)
    : PairPotential(sys, msg, nlist, val, p) {
  if (!nlist_)
    fail(msg_, "Lennard Jones pair potential requires neighbour list. None given.",
         "Neighbour list required by Lennard Jones potential, but non specified.");
  for (const char* k : {"epsilon", "sigma", "rcut", "use_particle_radii", "phase_in", "shifted", "WCA"})
    known_.push_back(k);
  std::string bad = params_ok(p);
  if (!bad.empty())
    fail(msg_, "Parameter \"" + bad + "\" is not a valid parameter for Lennard-Jones pair potential.",
         "Unknown parameter \"" + bad + "\" in Lennard-Jones potential.");
  if (has(p, "phase_in"))
    fail(msg_, "Lennard Jones pair potential: phase_in (population growth) is not supported by this implementation.",
         "Unsupported parameter phase_in.");
  if (has(p, "epsilon")) eps_ = to_double(p["epsilon"], "epsilon");
  else msg_->msg(Messenger::WARNING, "No potential depth (epsilon) specified for the Lennard Jones pair potential. Setting it to 1.");
  if (has(p, "sigma")) sigma_ = to_double(p["sigma"], "sigma");
  else msg_->msg(Messenger::WARNING, "No particle diameter (sigma) specified for the Lennard Jones pair potential. Setting it to 1.");
  if (has(p, "rcut")) rcut_ = to_double(p["rcut"], "rcut");
  else msg_->msg(Messenger::WARNING, "No cutoff distance (rcut) specified for the Lennard Jones pair potential. Setting it to 3.0.");
  bool shifted = false;
  if (has(p, "WCA")) { rcut_ = 1.122462048309373 * sigma_; shifted = true; }
  if (has(p, "shifted")) shifted = true;
  if (rcut_ > nlist_->get_cutoff())
    msg_->msg(Messenger::WARNING, "Neighbour list cutoff distance (" + std::to_string(nlist_->get_cutoff()) +
                                      " is smaller than the Lennard-Jones cuttof distance (" + std::to_string(rcut_) +
                                      "). Results will not be reliable.");
  sys_->check(samos_pair_lj(sys_->ctx(), eps_, sigma_, rcut_, shifted ? 1 : 0, has(p, "use_particle_radii") ? 1 : 0));
}

void PairLJPotential::set_pair_parameters(pairs_type& pp) {
  need_types(msg_, pp, "Lennard-Jones potential");
  const int t1 = to_int(pp["type_1"], "type_1"), t2 = to_int(pp["type_2"], "type_2");
  const double eps = has(pp, "epsilon") ? to_double(pp["epsilon"], "epsilon") : eps_;
  const double sig = has(pp, "sigma") ? to_double(pp["sigma"], "sigma") : sigma_;
  double rc = has(pp, "rcut") ? to_double(pp["rcut"], "rcut") : rcut_;
  int shifted = 0;
  if (has(pp, "WCA")) { rc = 1.122462048309373 * sig; shifted = 1; }
  sys_->check(samos_pair_type_param(sys_->ctx(), SAMOS_PAIR_LJ, t1, t2, eps, sig, rc, shifted));
}

void PairLJPotential::compute(double) { sys_->check(samos_pair_compute(sys_->ctx(), SAMOS_PAIR_LJ, &energy_)); }

// ---------------------------------------------------------------------------
// Aligners (src/aligner/pair/pair_polar_aligner.hpp:65-166,
// src/aligner/pair/pair_nematic_aligner.hpp:64-157)
// ---------------------------------------------------------------------------
static void read_J_rcut(const MessengerPtr& msg, pairs_type& p, const NeighbourListPtr& nl, double& J, double& rc,
                        const char* what) {
  if (has(p, "J")) J = to_double(p["J"], "J");
  else msg->msg(Messenger::WARNING, std::string("No coupling constant (J) specified for ") + what + ". Setting it to 1.");
  if (has(p, "rcut")) rc = to_double(p["rcut"], "rcut");
  else {
    msg->msg(Messenger::WARNING, std::string("No cutoff distance (rcut) specified for ") + what +
                                     ". Setting it to the global neighbour list cutoff.");
    rc = nl->get_cutoff();
  }
}

PairPolarAlign::PairPolarAlign(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, pairs_type& p)
    : PairAlign(sys, msg, nlist, p) {
  read_J_rcut(msg_, p, nlist_, J_, rcut_, "polar alignment");
  if (has(p, "exclude_boundary"))
    fail(msg_, "Polar alignment: exclude_boundary (tissue boundaries) is not supported by this implementation.",
         "Unsupported parameter exclude_boundary.");
  sys_->check(samos_align_polar(sys_->ctx(), J_, rcut_));
}

void PairPolarAlign::set_pair_parameters(pairs_type& pp) {
  if (!has(pp, "type_1"))
    fail(msg_, "type_1 has not been defined for pairwise alignment in polar aligner.", "Missing key for pair alignment parameters.");
  if (!has(pp, "type_2"))
    fail(msg_, "type_2 has not been defined for pairwise alignment in polar aligner.", "Missing key for pair potential parameters.");
  const double J = has(pp, "J") ? to_double(pp["J"], "J") : J_;
  const double rc = has(pp, "rcut") ? to_double(pp["rcut"], "rcut") : rcut_;
  sys_->check(samos_align_type_param(sys_->ctx(), SAMOS_ALIGN_POLAR, to_int(pp["type_1"], "type_1"),
                                     to_int(pp["type_2"], "type_2"), J, rc));
}

void PairPolarAlign::compute() { sys_->check(samos_align_compute(sys_->ctx(), SAMOS_ALIGN_POLAR, &energy_)); }

PairNematicAlign::PairNematicAlign(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, pairs_type& p)
    : PairAlign(sys, msg, nlist, p) {
  read_J_rcut(msg_, p, nlist_, J_, rcut_, "MF alignment");
  sys_->check(samos_align_nematic(sys_->ctx(), J_, rcut_));
}

void PairNematicAlign::set_pair_parameters(pairs_type& pp) {
  if (!has(pp, "type_1"))
    fail(msg_, "type_1 has not been defined for pairwise alignment in MF aligner.", "Missing key for pair alignment parameters.");
  if (!has(pp, "type_2"))
    fail(msg_, "type_2 has not been defined for pairwise alignment in MF aligner.", "Missing key for pair potential parameters.");
  const double J = has(pp, "J") ? to_double(pp["J"], "J") : J_;
  const double rc = has(pp, "rcut") ? to_double(pp["rcut"], "rcut") : rcut_;
  sys_->check(samos_align_type_param(sys_->ctx(), SAMOS_ALIGN_NEMATIC, to_int(pp["type_1"], "type_1"),
                                     to_int(pp["type_2"], "type_2"), J, rc));
}

void PairNematicAlign::compute() { sys_->check(samos_align_compute(sys_->ctx(), SAMOS_ALIGN_NEMATIC, &energy_)); }

// ---------------------------------------------------------------------------
// Constraints (src/constraints/constraint.hpp:56-120 and the per-surface headers)
// ---------------------------------------------------------------------------
Constraint::Constraint(SystemPtr sys, MessengerPtr msg, pairs_type& p) : sys_(sys), msg_(msg) {
  // rescale, rescale_steps, rescale_freq (constraint.hpp:85-108); the steps and
  // frequency are read as numbers and stored as ints, as there
  if (has(p, "rescale")) {
    rescale_ = to_double(p["rescale"], "rescale");
    msg_->msg(Messenger::INFO, "Constraint. Rescaling constraint by factor " + p["rescale"] + ".");
  }
  if (has(p, "rescale_steps")) rescale_steps_ = (int)to_double(p["rescale_steps"], "rescale_steps");
  if (has(p, "rescale_freq")) rescale_freq_ = (int)to_double(p["rescale_freq"], "rescale_freq");
  if (has(p, "group") && p["group"] != "all")
    fail(msg_, "Constraints on groups other than \"all\" are not supported by this implementation.", "Unsupported group.");
  // The reference reads maxiter and tol for every constraint; both go through an
  // integer conversion (constraint.hpp:71,82), so a fractional tolerance such as
  // 1e-8 is a conversion error there as here.
  if (has(p, "maxiter")) {
    max_iter_ = to_int(p["maxiter"], "maxiter");
    msg_->msg(Messenger::INFO, "Constraint. Maximum number of iterations set to " + p["maxiter"] + ".");
  } else {
    msg_->msg(Messenger::WARNING, "Constraint. Maximum number of iterations has not been set. Assuming 100");
  }
  if (has(p, "tol")) {
    tol_ = (double)to_int(p["tol"], "tol");
    msg_->msg(Messenger::INFO, "Constraint. Tolerance set to " + p["tol"] + ".");
  } else {
    msg_->msg(Messenger::WARNING, "Constraint. Tolerance has not been set. Assuming 1e-6.");
  }
}

ConstraintNone::ConstraintNone(SystemPtr sys, MessengerPtr msg, pairs_type& p) : Constraint(sys, msg, p) {
  msg_->msg(Messenger::INFO, "Constraint none.");
}

ConstraintPlane::ConstraintPlane(SystemPtr sys, MessengerPtr msg, pairs_type& p) : Constraint(sys, msg, p) {
  unlimited_ = has(p, "unlimited");
  if (has(p, "zpos")) zpos_ = to_double(p["zpos"], "zpos");
  else msg_->msg(Messenger::WARNING, "Plane constraint. No zpos specified. Assuming that the constraint plane is at z = 0.");
}

ConstraintSphere::ConstraintSphere(SystemPtr sys, MessengerPtr msg, pairs_type& p) : Constraint(sys, msg, p) {
  if (has(p, "r")) r_ = to_double(p["r"], "r");
  else msg_->msg(Messenger::WARNING, "Spherical constraint. No radius set. Assuming 1.");
}

ConstraintSurface::ConstraintSurface(SystemPtr sys, MessengerPtr msg, pairs_type& p, const std::string& name)
    : Constraint(sys, msg, p) {
  struct Spec { int kind; std::vector<std::pair<const char*, double>> params; std::vector<std::string> warn; };
  static const std::map<std::string, Spec> specs = {
      {"cylinder", {SAMOS_CONSTRAINT_CYLINDER, {{"r", 1.0}}, {"Cylindrical constraint. No radius set. Assuming 1."}}},
      {"torus", {SAMOS_CONSTRAINT_TORUS, {{"c", 1.0}, {"a", 0.5}},
                 {"Toroidal constraint. No torus radius set. Assuming 1.",
                  "Toroidal constraint. No tube radius set. Assuming 0.5."}}},
      {"ellipsoid", {SAMOS_CONSTRAINT_ELLIPSOID, {{"a", 1.0}, {"b", 1.0}, {"c", 1.0}},
                     {"Ellipsoid constraint. No parameter a. Assuming 1.", "Ellipsoid constraint. No parameter b. Assuming 1.",
                      "Ellipsoid constraint. No parameter c. Assuming 1."}}}};
  const Spec& sp = specs.at(name);
  kind_ = sp.kind;
  for (size_t i = 0; i < sp.params.size(); ++i) {
    const char* key = sp.params[i].first;
    if (has(p, key)) q_.push_back(to_double(p[key], key));
    else {
      msg_->msg(Messenger::WARNING, sp.warn[i]);
      q_.push_back(sp.params[i].second);
    }
  }
  if (kind_ == SAMOS_CONSTRAINT_TORUS && q_[1] >= q_[0])
    fail(msg_, "Toroidal constraint. Tube radius has to be smaller than torus radius.",
         "Toroidal constraint. Incompatible parameters.");
}

// ---------------------------------------------------------------------------
// Integrators (src/integrators/integrator.hpp:67-123 and the subclasses)
// ---------------------------------------------------------------------------
Integrator::Integrator(SystemPtr sys, MessengerPtr msg, ValuePtr temp, pairs_type& p)
    : sys_(sys), msg_(msg), temp_(temp) {
  known_ = {"dt", "group", "temperature_control", "min_val", "max_val", "steps"};
  if (has(p, "dt")) {
    msg_->msg(Messenger::WARNING, "Integrator is setting its own time step size. While this is allowed, it may cause "
                                  "odd behaviour in case different integrators set different internal step size.");
    dt_ = to_double(p["dt"], "dt");
  } else {
    dt_ = sys_->integrator_step();
  }
  if (dt_ <= 0.0)
    fail(msg_, "Integrator step size has to be larger than 0. Did you forget to use \"timestep\" command?",
         "Integrator step has not been set.");
  if (has(p, "group") && p["group"] != "all") {
    const std::string g = p["group"];
    if (!sys_->has_group(g))
      fail(msg_, "Group " + g + " does not exist.", "Integrator group does not exist.");
    msg_->msg(Messenger::INFO, "Applying integrator to group " + g + ".");
    sys_->check(samos_set_integrator_group(sys_->ctx(), g.c_str()));
  }
}

std::string Integrator::params_ok(pairs_type& params) const {
  for (auto& kv : params)
    if (std::find(known_.begin(), known_.end(), kv.first) == known_.end()) return kv.first;
  return "";
}

void Integrator::check_params(pairs_type& p, const std::string& what) {
  std::string bad = params_ok(p);
  if (!bad.empty())
    fail(msg_, "Parameter \"" + bad + "\" is not a valid parameter for " + what + " integrator.",
         "Unknown parameter \"" + bad + "\" in " + what + " integrator.");
}

// RAII swap of the context's step size while an integrator with its own dt registers.
struct StepScope {
  SystemPtr s;
  double keep;
  StepScope(SystemPtr s_, double dt) : s(s_), keep(s_->integrator_step()) { s->check(samos_set_timestep(s->ctx(), dt)); }
  ~StepScope() { samos_set_timestep(s->ctx(), keep); }
};

static double getd(pairs_type& p, const char* k, double dflt) { return has(p, k) ? to_double(p[k], k) : dflt; }

// which: 0 brownian (integrator_brownian.hpp:55-160), 1 brownian_pos, 2 brownian_align
IntegratorBrownian::IntegratorBrownian(SystemPtr sys, MessengerPtr msg, ValuePtr temp, pairs_type& p, int which)
    : Integrator(sys, msg, temp, p) {
  static const std::vector<std::vector<std::string>> extra = {
      {"v0", "nu", "mu", "mur", "seed", "nematic", "tau", "velocity_align"},
      {"mu", "seed"},
      {"nu", "mur", "seed", "nematic", "tau"}};
  for (auto& k : extra[which]) known_.push_back(k);
  static const char* names[] = {"brownian", "brownian_pos", "brownian_align"};
  check_params(p, names[which]);
  const double v0 = getd(p, "v0", 1.0), nu = getd(p, "nu", 1.0), mu = getd(p, "mu", 1.0), mur = getd(p, "mur", 1.0);
  const int seed = has(p, "seed") ? to_int(p["seed"], "seed") : 0;
  const bool nematic = has(p, "nematic");
  const double tau = getd(p, "tau", 1.0);
  StepScope sc(sys_, dt_);
  sys_->check(samos_integrator_brownian(sys_->ctx(), which, v0, nu, mu, mur, seed, nematic ? 1 : 0, tau,
                                        has(p, "velocity_align") ? 1 : 0, temp_->kind, temp_->min_val,
                                        temp_->max_val, temp_->steps));
}

// IntegratorLangevin (src/integrators/integrator_langevin.hpp:64-121)
IntegratorLangevin::IntegratorLangevin(SystemPtr sys, MessengerPtr msg, ValuePtr temp, pairs_type& p)
    : Integrator(sys, msg, temp, p) {
  for (const char* k : {"gamma", "seed", "method"}) known_.push_back(k);
  check_params(p, "langevin");
  const double gamma = getd(p, "gamma", 1.0);
  const int seed = has(p, "seed") ? to_int(p["seed"], "seed") : 0;
  std::string m = has(p, "method") ? p["method"] : "baoab";
  std::transform(m.begin(), m.end(), m.begin(), [](unsigned char c) { return std::tolower(c); });
  const int method = m == "baoab" ? SAMOS_LANGEVIN_BAOAB : m == "spv" ? SAMOS_LANGEVIN_SPV : m == "bbk" ? SAMOS_LANGEVIN_BBK : -1;
  if (method < 0) fail(msg_, "Unknown integration method " + m + ".", "Unknown integration method");
  StepScope sc(sys_, dt_);
  sys_->check(samos_integrator_langevin(sys_->ctx(), gamma, seed, method, temp_->kind, temp_->min_val, temp_->max_val,
                                        temp_->steps));
}

// IntegratorFIRE (src/integrators/integrator_fire.hpp:58-154): no unknown-parameter check.
IntegratorFIRE::IntegratorFIRE(SystemPtr sys, MessengerPtr msg, ValuePtr temp, pairs_type& p)
    : Integrator(sys, msg, temp, p) {
  StepScope sc(sys_, dt_);
  sys_->check(samos_integrator_fire(sys_->ctx(), getd(p, "alpha", 0.1), getd(p, "alpha_scale", 0.99),
                                    getd(p, "f_tolerance", 1e-4), getd(p, "e_tolerance", 1e-6),
                                    getd(p, "dt_inc", 1.1), getd(p, "dt_dec", 0.5), getd(p, "dt_max", 1.0),
                                    has(p, "min_neg_steps") ? to_int(p["min_neg_steps"], "min_neg_steps") : 5));
}

// IntegratorNVE (src/integrators/integrator_nve.hpp:59-77): `limit` and
// `angle_limit` are optional; no unknown-parameter check.
IntegratorNVE::IntegratorNVE(SystemPtr sys, MessengerPtr msg, ValuePtr temp, pairs_type& p)
    : Integrator(sys, msg, temp, p) {
  double limit = -1.0, angle = -1.0;
  if (has(p, "limit")) {
    msg_->msg(Messenger::WARNING, "NVE integrator has limit on the maximum particle displacement. It is set to " +
                                      p["limit"] + ".");
    limit = to_double(p["limit"], "limit");
  }
  if (has(p, "angle_limit")) {
    msg_->msg(Messenger::WARNING, "NVE integrator has limit on the maximum rotation of the director. It is set to " +
                                      p["angle_limit"] + ".");
    angle = to_double(p["angle_limit"], "angle_limit");
  }
  StepScope sc(sys_, dt_);
  sys_->check(samos_integrator_nve(sys_->ctx(), limit, angle));
}

// ---------------------------------------------------------------------------
// Registries (src/utils/register_pair_potentials.cpp:35,39,
// register_pair_aligners.cpp:35,37, register_integrators.cpp:35-53,
// register_constraints.cpp:35,37,57)
// ---------------------------------------------------------------------------
std::map<std::string, PairPotentialFactory>& pair_potentials() {
  static std::map<std::string, PairPotentialFactory> m = {
      {"soft", [](SystemPtr s, MessengerPtr m_, NeighbourListPtr n, ValuePtr v, pairs_type& p) -> PairPotentialPtr {
         return std::make_shared<PairSoftPotential>(s, m_, n, v, p);
       }},
      {"lj", [](SystemPtr s, MessengerPtr m_, NeighbourListPtr n, ValuePtr v, pairs_type& p) -> PairPotentialPtr {
         return std::make_shared<PairLJPotential>(s, m_, n, v, p);
       }}};
  static bool extras_added = false;
  if (!extras_added) {
    extras_added = true;
    for (const char* nm : {"morse", "gaussian", "yukawa", "soft_attractive"}) {
      const std::string name = nm;
      m[name] = [name](SystemPtr s, MessengerPtr m_, NeighbourListPtr n, ValuePtr v, pairs_type& p) -> PairPotentialPtr {
        return std::make_shared<PairExtraPotential>(s, m_, n, v, p, name);
      };
    }
  }
  return m;
}

std::map<std::string, PairAlignFactory>& pair_aligners() {
  static std::map<std::string, PairAlignFactory> m = {
      {"polar", [](SystemPtr s, MessengerPtr m_, NeighbourListPtr n, pairs_type& p) -> PairAlignPtr {
         return std::make_shared<PairPolarAlign>(s, m_, n, p);
       }},
      {"nematic", [](SystemPtr s, MessengerPtr m_, NeighbourListPtr n, pairs_type& p) -> PairAlignPtr {
         return std::make_shared<PairNematicAlign>(s, m_, n, p);
       }}};
  return m;
}

std::map<std::string, ConstraintFactory>& constraints() {
  static std::map<std::string, ConstraintFactory> m = {
      {"none", [](SystemPtr s, MessengerPtr m_, pairs_type& p) -> ConstraintPtr { return std::make_shared<ConstraintNone>(s, m_, p); }},
      {"plane", [](SystemPtr s, MessengerPtr m_, pairs_type& p) -> ConstraintPtr { return std::make_shared<ConstraintPlane>(s, m_, p); }},
      {"sphere", [](SystemPtr s, MessengerPtr m_, pairs_type& p) -> ConstraintPtr { return std::make_shared<ConstraintSphere>(s, m_, p); }},
      {"cylinder", [](SystemPtr s, MessengerPtr m_, pairs_type& p) -> ConstraintPtr { return std::make_shared<ConstraintSurface>(s, m_, p, "cylinder"); }},
      {"torus", [](SystemPtr s, MessengerPtr m_, pairs_type& p) -> ConstraintPtr { return std::make_shared<ConstraintSurface>(s, m_, p, "torus"); }},
      {"ellipsoid", [](SystemPtr s, MessengerPtr m_, pairs_type& p) -> ConstraintPtr { return std::make_shared<ConstraintSurface>(s, m_, p, "ellipsoid"); }}};
  return m;
}

std::map<std::string, IntegratorFactory>& integrators() {
  static std::map<std::string, IntegratorFactory> m = {
      {"brownian", [](SystemPtr s, MessengerPtr m_, ValuePtr t, pairs_type& p) -> IntegratorPtr { return std::make_shared<IntegratorBrownian>(s, m_, t, p, 0); }},
      {"brownian_pos", [](SystemPtr s, MessengerPtr m_, ValuePtr t, pairs_type& p) -> IntegratorPtr { return std::make_shared<IntegratorBrownian>(s, m_, t, p, 1); }},
      {"brownian_align", [](SystemPtr s, MessengerPtr m_, ValuePtr t, pairs_type& p) -> IntegratorPtr { return std::make_shared<IntegratorBrownian>(s, m_, t, p, 2); }},
      {"langevin", [](SystemPtr s, MessengerPtr m_, ValuePtr t, pairs_type& p) -> IntegratorPtr { return std::make_shared<IntegratorLangevin>(s, m_, t, p); }},
      {"fire", [](SystemPtr s, MessengerPtr m_, ValuePtr t, pairs_type& p) -> IntegratorPtr { return std::make_shared<IntegratorFIRE>(s, m_, t, p); }},
      {"nve", [](SystemPtr s, MessengerPtr m_, ValuePtr t, pairs_type& p) -> IntegratorPtr { return std::make_shared<IntegratorNVE>(s, m_, t, p); }}};
  return m;
}

// ---------------------------------------------------------------------------
// Logger (src/log/logger.cpp:40-137, src/log/log_*.hpp)
// ---------------------------------------------------------------------------
static const char* kLogQuantities[] = {"velocity", "vec_velocity", "soft_energy", "lj_energy",
                                       "polar_align", "nematic_align", "kinetic_energy"};

Logger::Logger(SystemPtr sys, MessengerPtr msg, const std::string& file, pairs_type& params)
    : sys_(sys), msg_(msg), out_(file) {
  quantities_.push_back("step");
  if (has(params, "freq")) freq_ = to_int(params["freq"], "freq");
  else msg_->msg(Messenger::WARNING, "No log frequency specified. Using default of logging each 100 time steps.");
  for (auto& kv : params) {
    std::string q = kv.first;
    std::transform(q.begin(), q.end(), q.begin(), [](unsigned char c) { return std::tolower(c); });
    if (q == "freq") continue;
    bool ok = std::find(std::begin(kLogQuantities), std::end(kLogQuantities), q) != std::end(kLogQuantities);
    if (!ok) fail(msg_, "Unknown log quantity : " + q + ".", "Unknown log quantity.");
    quantities_.push_back(q);
  }
  std::time_t now = std::time(nullptr);
  char buf[64];
  std::strftime(buf, sizeof buf, "%Y-%b-%d %H:%M:%S", std::localtime(&now));
  out_ << "# Log file : " << file << "\n# Generated on : " << buf << "\n# ";
  for (size_t i = 0; i < quantities_.size(); ++i) out_ << "(" << i + 1 << ") " << quantities_[i] << " ";
  out_ << std::endl;
}

void Logger::log(int step) {
  if (!due(step)) return;
  double e[4] = {0, 0, 0, 0};
  bool have_e = false;
  char buf[64];
  for (auto& q : quantities_) {
    if (q == "step") { std::snprintf(buf, sizeof buf, "%8d ", step); out_ << buf << " "; continue; }
    double v = 0.0;
    if (q == "velocity") { sys_->check(samos_reduce(sys_->ctx(), 4, &v)); }
    else if (q == "vec_velocity") { sys_->check(samos_reduce(sys_->ctx(), 3, &v)); }
    else if (q == "kinetic_energy") { sys_->check(samos_reduce(sys_->ctx(), 2, &v)); v /= sys_->size(); }
    else {
      if (!have_e) { sys_->check(samos_energies(sys_->ctx(), e)); have_e = true; }
      v = q == "soft_energy" ? e[0] : q == "lj_energy" ? e[1] : q == "polar_align" ? e[2] : e[3];
    }
    std::snprintf(buf, sizeof buf, "%12.6e ", v);
    out_ << buf << " ";
  }
  out_ << std::endl;
}

// ---------------------------------------------------------------------------
// Dump (src/dump/dump.cpp:60-413, 419-800): xyz, full, input, velocity,
// director, xyzv and dcd formats; vtp is accepted and writes nothing, as in a
// reference build without VTK (dump.cpp:403-406).
// ---------------------------------------------------------------------------
Dump::Dump(SystemPtr sys, MessengerPtr msg, const std::string& name, pairs_type& params)
    : sys_(sys), msg_(msg), name_(name), params_(params) {
  if (!has(params, "type")) fail(msg_, "No dump type specified.", "No dump type specified.");
  type_ = params["type"];
  static const std::map<std::string, std::string> ext = {{"xyz", "xyz"},      {"full", "dat"}, {"input", "inp"},
                                                         {"velocity", "vel"}, {"dcd", "dcd"},  {"director", "dir"},
                                                         {"xyzv", "xyzv"},    {"vtp", "vtp"}};
  auto it = ext.find(type_);
  if (it == ext.end())
    fail(msg_, "Dump type " + type_ + " is not supported by this implementation (xyz, full, input, velocity, director, xyzv, "
                                      "dcd, vtp).", "Unsupported dump type");
  ext_ = it->second;
  if (has(params, "directory")) dir_ = params["directory"];
  if (has(params, "start")) start_ = to_int(params["start"], "start");
  freq_ = has(params, "freq") ? to_int(params["freq"], "freq") : 100;
  if (has(params, "scale")) scale_ = to_double(params["scale"], "scale");
  multi_ = has(params, "multi");
  header_ = has(params, "header");
  keys_ = has(params, "keys");
  if (has(params, "group")) group_ = params["group"];
  if (!sys_->has_group(group_))
    fail(msg_, "Dump: Group " + group_ + " is not defined.", "Attempted to dump group that does not exist.");
  if (type_ == "input") {
    for (const char* k : {"id", "tp", "radius", "coordinate", "velocity", "director", "omega"}) params_[k] = "";
    keys_ = true;
  }
  if (type_ == "dcd") {
    if (group_ != "all") fail(msg_, "DCD files does not group specific dump.", "Attempted to dump groups into a DCD output.");
    if (multi_)
      fail(msg_, "DCD files cannot be split. mulifile flag is not permitted.",
           "Attempted to spread DCD output into multiple files.");
    single_.open(dir_ + "/" + name_ + "." + ext_, std::ios::binary);
    write_dcd_header();
    return;
  }
  if (type_ == "vtp") {
    msg_->msg(Messenger::WARNING, "vtp dumps need VTK, which this implementation does not use: no file is written.");
    return;
  }
  if (!multi_) single_.open(dir_ + "/" + name_ + "." + ext_);
}

namespace {
void put_int(std::ostream& o, int v) { o.write(reinterpret_cast<const char*>(&v), 4); }
}  // namespace

// The CHARMM-style header of the reference (dump.cpp:296-341): 84-byte block
// with "CORD", start and frequency, unit-cell flag and version 24; two
// 80-byte title lines; the particle count.
void Dump::write_dcd_header() {
  std::ostream& f = single_;
  put_int(f, 84);
  f.write("CORD", 4);
  put_int(f, 0);
  put_int(f, start_);
  put_int(f, freq_);
  for (int k = 0; k < 6; ++k) put_int(f, 0);
  put_int(f, 0);
  put_int(f, 1);
  for (int k = 0; k < 8; ++k) put_int(f, 0);
  put_int(f, 24);
  put_int(f, 84);
  put_int(f, 164);
  put_int(f, 2);
  char title[80] = {0};
  std::snprintf(title, sizeof title, "Created by SAMoS");
  f.write(title, 80);
  char when[81] = {0};
  const std::time_t now = std::time(nullptr);
  std::strftime(when, 80, "REMARKS Created  %d %B, %Y at %H:%M", std::localtime(&now));
  f.write(when, 80);
  put_int(f, 164);
  put_int(f, 4);
  put_int(f, sys_->size());
  put_int(f, 4);
  if (!f.good()) fail(msg_, "Error writing DCD header", "Error writing DCD file");
}

// One frame (dump.cpp:419-488): the unit-cell record (six doubles, 100 on the
// diagonal), then x, y and z as FP32 records, unwrapped with `unwrap`.
void Dump::write_dcd_frame() {
  std::ostream& f = single_;
  const int n = sys_->size();
  const double cell[6] = {100.0, 0.0, 100.0, 0.0, 0.0, 100.0};
  put_int(f, 48);
  f.write(reinterpret_cast<const char*>(cell), 48);
  put_int(f, 48);
  const bool unwrap = params_.count("unwrap") != 0;
  const Box& b = sys_->box();
  const int fc[3] = {SAMOS_X, SAMOS_Y, SAMOS_Z}, ic[3] = {SAMOS_IX, SAMOS_IY, SAMOS_IZ};
  const double L[3] = {b.Lx, b.Ly, b.Lz};
  std::vector<float> buf(n);
  for (int k = 0; k < 3; ++k) {
    const std::vector<double> c = sys_->column(fc[k]);
    std::vector<int> im;
    if (unwrap) im = sys_->icolumn(ic[k]);
    for (int i = 0; i < n; ++i) buf[i] = static_cast<float>(unwrap ? c[i] + im[i] * L[k] : c[i]);
    put_int(f, n * 4);
    f.write(reinterpret_cast<const char*>(buf.data()), 4L * n);
    put_int(f, n * 4);
  }
  if (!f.good()) fail(msg_, "Error writing into DCD file.", "Error writing DCD file");
}

void Dump::write(std::ostream& o) {
  const std::vector<int>& ids = sys_->group(group_);
  const int n = (int)ids.size();
  char buf[256];
  if (type_ == "xyz") {
    auto x = sys_->column(SAMOS_X), y = sys_->column(SAMOS_Y), z = sys_->column(SAMOS_Z);
    auto t = sys_->icolumn(SAMOS_TYPE);
    o << n << "\nGenerated by SAMoS code.\n";
    for (int i : ids) {
      std::snprintf(buf, sizeof buf, "%3d\t%10.6f\t%10.6f\t%10.6f", t[i], x[i], y[i], z[i]);
      o << buf << "\n";
    }
    return;
  }
  if (type_ == "velocity" || type_ == "director" || type_ == "xyzv") {
    // dump.cpp:710-782: vector columns scaled by `scale`
    const bool vel = type_ == "velocity" || (type_ == "xyzv" && params_.count("velocity"));
    auto x = sys_->column(SAMOS_X), y = sys_->column(SAMOS_Y), z = sys_->column(SAMOS_Z);
    auto ax = sys_->column(vel ? SAMOS_VX : SAMOS_NX), ay = sys_->column(vel ? SAMOS_VY : SAMOS_NY),
         az = sys_->column(vel ? SAMOS_VZ : SAMOS_NZ);
    const double a = scale_;
    if (type_ == "xyzv") {
      auto t = sys_->icolumn(SAMOS_TYPE);
      o << n << "\n";
      o << "Generated by SAMoS code. Vectors are " << (vel ? "velocities" : "directors") << " scaled by " << a << "\n";
      for (int i : ids) {
        std::snprintf(buf, sizeof buf, "%c\t%12.7e\t%12.7e\t%12.7e\t%12.7e\t%12.7e\t%12.7e", (char)(64 + t[i]), x[i],
                      y[i], z[i], a * ax[i], a * ay[i], a * az[i]);
        o << buf << "\n";
      }
      return;
    }
    if (header_) {
      const char c = vel ? 'v' : 'n';
      o << "# scale a = " << a << "\n";
      o << "# x-0.5*a*" << c << "x  y-0.5*a*" << c << "y z-0.5*a*" << c << "z  a*" << c << "x  a*" << c << "y  a*" << c
        << "z\n";
    }
    for (int i : ids) {
      std::snprintf(buf, sizeof buf, "%12.7e\t%12.7e\t%12.7e\t%12.7e\t%12.7e\t%12.7e", x[i] - 0.5 * a * ax[i],
                    y[i] - 0.5 * a * ay[i], z[i] - 0.5 * a * az[i], a * ax[i], a * ay[i], a * az[i]);
      o << buf << "\n";
    }
    return;
  }
  auto on = [&](const char* k) { return params_.find(k) != params_.end(); };
  const char* order[] = {"id", "tp", "radius", "coordinate", "velocity", "force", "director", "omega", "image_flags"};
  const char* heads[] = {" id ", " type ", " radius ", " x  y  z ", " vx  vy  vz ", " fx  fy  fz ", " nx  ny  nz ",
                         " omega ", " ix  iy  iz "};
  if (header_) {
    o << "# ";
    for (int k = 0; k < 9; ++k) if (on(order[k])) o << heads[k];
    o << "\n";
  }
  if (keys_) {
    o << "keys: ";
    for (int k = 0; k < 9; ++k) if (on(order[k])) o << heads[k];
    o << "\n";
  }
  std::vector<double> x, y, z, vx, vy, vz, fx, fy, fz, nx, ny, nz, om, r;
  std::vector<int> tp, ix, iy, iz;
  if (on("tp")) tp = sys_->icolumn(SAMOS_TYPE);
  if (on("radius")) r = sys_->column(SAMOS_RADIUS);
  if (on("coordinate")) { x = sys_->column(SAMOS_X); y = sys_->column(SAMOS_Y); z = sys_->column(SAMOS_Z); }
  if (on("coordinate") && on("unwrap")) { ix = sys_->icolumn(SAMOS_IX); iy = sys_->icolumn(SAMOS_IY); iz = sys_->icolumn(SAMOS_IZ); }
  if (on("velocity")) { vx = sys_->column(SAMOS_VX); vy = sys_->column(SAMOS_VY); vz = sys_->column(SAMOS_VZ); }
  if (on("force")) { fx = sys_->column(SAMOS_FX); fy = sys_->column(SAMOS_FY); fz = sys_->column(SAMOS_FZ); }
  if (on("director")) { nx = sys_->column(SAMOS_NX); ny = sys_->column(SAMOS_NY); nz = sys_->column(SAMOS_NZ); }
  if (on("omega")) om = sys_->column(SAMOS_OMEGA);
  if (on("image_flags") && ix.empty()) { ix = sys_->icolumn(SAMOS_IX); iy = sys_->icolumn(SAMOS_IY); iz = sys_->icolumn(SAMOS_IZ); }
  const Box& b = sys_->box();
  for (int i : ids) {
    if (on("id")) { std::snprintf(buf, sizeof buf, "%5d ", i); o << buf; }
    if (on("tp")) { std::snprintf(buf, sizeof buf, "%2d ", tp[i]); o << buf; }
    if (on("radius")) { std::snprintf(buf, sizeof buf, "%8.5f ", r[i]); o << buf; }
    if (on("coordinate")) {
      if (on("unwrap"))
        std::snprintf(buf, sizeof buf, " %16.10e  %16.10e  %16.10e", x[i] + ix[i] * b.Lx, y[i] + iy[i] * b.Ly, z[i] + iz[i] * b.Lz);
      else
        std::snprintf(buf, sizeof buf, " %16.10e  %16.10e  %16.10e", x[i], y[i], z[i]);
      o << buf;
    }
    if (on("velocity")) { std::snprintf(buf, sizeof buf, " %12.7e  %12.7e  %12.7e", vx[i], vy[i], vz[i]); o << buf; }
    if (on("force")) { std::snprintf(buf, sizeof buf, " %12.7e  %12.7e  %12.7e", fx[i], fy[i], fz[i]); o << buf; }
    if (on("director")) { std::snprintf(buf, sizeof buf, " %10.6f  %10.6f  %10.6f", nx[i], ny[i], nz[i]); o << buf; }
    if (on("omega")) { std::snprintf(buf, sizeof buf, "%8.5f ", om[i]); o << buf; }
    if (on("image_flags")) { std::snprintf(buf, sizeof buf, " %3d  %3d  %3d ", ix[i], iy[i], iz[i]); o << buf; }
    o << "\n";
  }
}

void Dump::dump(int step) {
  if (!due(step) || type_ == "vtp") return;
  if (type_ == "dcd") {
    write_dcd_frame();
    single_.flush();
    return;
  }
  if (multi_) {
    char num[32];
    std::snprintf(num, sizeof num, "%010d", step);
    std::ofstream f(dir_ + "/" + name_ + "_" + num + "." + ext_);
    write(f);
  } else {
    write(single_);
    single_.flush();
  }
}

// Further pair potentials with their reference parameter names and defaults:
//   morse            D 1, a 1, re 1, rcut 2.5 re     (pair_morse_potential.hpp:69-160)
//   gaussian         A 2, B -1, alpha 1, beta 3, rA 0, rB 1.5, rcut 3
//                    (pair_gaussian_potential.hpp:72-200; no unknown-parameter check)
//   yukawa           g 1, kappa 1, rcut 3            (pair_yukawa_potential.hpp:65-120)
//   soft_attractive  k 1, a 2, re_fact 1.25          (pair_soft_attractive_potential.hpp:70-140)
PairExtraPotential::PairExtraPotential(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, ValuePtr val,
                                       pairs_type& p, const std::string& name)
    : PairPotential(sys, msg, nlist, val, p), name_(name) {
  struct Spec { int kind; std::vector<std::pair<const char*, double>> params; bool check, shift; const char* label; };
  static const std::map<std::string, Spec> specs = {
      {"morse", {SAMOS_PAIR_MORSE, {{"D", 1.0}, {"a", 1.0}, {"re", 1.0}, {"rcut", -1.0}}, true, true, "Morse"}},
      {"gaussian", {SAMOS_PAIR_GAUSSIAN, {{"A", 2.0}, {"B", -1.0}, {"alpha", 1.0}, {"beta", 3.0}, {"rA", 0.0},
                                          {"rB", 1.5}, {"rcut", 3.0}}, false, true, "Gaussian"}},
      {"yukawa", {SAMOS_PAIR_YUKAWA, {{"g", 1.0}, {"kappa", 1.0}, {"rcut", 3.0}}, true, false, "Yukawa"}},
      {"soft_attractive", {SAMOS_PAIR_SOFT_ATTRACTIVE, {{"k", 1.0}, {"a", 2.0}, {"re_fact", 1.25}}, true, false,
                           "soft attractive"}}};
  if (!nlist_)
    fail(msg_, std::string(specs.at(name).label) + " pair potential requires neighbour list. None given.",
         "Neighbour list required by " + std::string(specs.at(name).label) + " potential, but non specified.");
  const Spec& sp = specs.at(name);
  for (auto& kv : sp.params) known_.push_back(kv.first);
  if (name != "yukawa") known_.push_back("use_particle_radii");
  if (name != "yukawa") known_.push_back("phase_in");
  if (sp.shift) known_.push_back("shifted");
  if (sp.check) {
    std::string bad = params_ok(p);
    if (!bad.empty())
      fail(msg_, "Parameter \"" + bad + "\" is not a valid parameter for " + sp.label + " pair potential.",
           "Unknown parameter \"" + bad + "\" in " + sp.label + " potential.");
  }
  if (has(p, "phase_in"))
    fail(msg_, std::string(sp.label) + " pair potential: phase_in (population growth) is not supported by this implementation.",
         "Unsupported parameter phase_in.");
  std::vector<double> q;
  for (auto& kv : sp.params) {
    double v = kv.second;
    if (has(p, kv.first)) v = to_double(p[kv.first], kv.first);
    else msg_->msg(Messenger::WARNING, std::string("No ") + kv.first + " specified for " + sp.label +
                                          " pair potential. Using the default.");
    q.push_back(v);
  }
  if (name == "morse" && !has(p, "rcut")) q[3] = 2.5 * q[2];
  const bool radii = has(p, "use_particle_radii") && name != "yukawa";
  const bool shifted = has(p, "shifted") && sp.shift;
  sys_->check(samos_pair_extra(sys_->ctx(), sp.kind, q.data(), (int)q.size(), radii ? 1 : 0, shifted ? 1 : 0));
}

void PairExtraPotential::set_pair_parameters(pairs_type&) {
  fail(msg_, "Per-type parameters of the " + name_ + " pair potential are not supported by this implementation.",
       "Unsupported pair_param for " + name_ + ".");
}

void PairExtraPotential::compute(double) { sys_->check(samos_pair_extra_energy(sys_->ctx(), &energy_)); }


}  // namespace samos_host
