// Host side of the SAMoS particle pipeline: the reference's plugin interfaces
// (PairPotential, PairAlign, Integrator, Constraint, Value), their name->factory
// registries, the configuration-file grammar and the observers (logs, dumps).
// Every object forwards its work to the device through the C-ABI of
// libsamos_hip.so (include/samos_hip.h); nothing here computes particle data.
#pragma once

#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "samos_hip.h"

namespace samos_host {

// Key/value pairs of one `{ ... }` block (src/parser/parse_parameters.hpp:43).
using pairs_type = std::map<std::string, std::string>;

// ---------------------------------------------------------------------------
// Messenger (src/messenger/messenger.cpp:33-75): INFO / WARNING / ERROR lines.
// ---------------------------------------------------------------------------
class Messenger {
 public:
  enum Level { INFO, WARNING, ERROR };
  explicit Messenger(const std::string& file = "");
  void msg(Level l, const std::string& text);
  void set_file(const std::string& file);

 private:
  std::ofstream out_;
  bool to_file_ = false;
};
using MessengerPtr = std::shared_ptr<Messenger>;

// ---------------------------------------------------------------------------
// Configuration grammar (src/parser/parse_command.hpp:61-110,
// src/parser/parse_parameters.hpp:45-59).
// ---------------------------------------------------------------------------
struct CommandLine {
  std::string command;
  std::string rest;  // attribute and parameter block, unparsed
};
bool parse_command(const std::string& line, CommandLine& out);
// `attr { k = v; flag; ... }`: attribute is the first word before '{' (may be empty).
bool parse_attrib_params(const std::string& rest, std::string& attrib, pairs_type& params);
bool parse_params(const std::string& text, pairs_type& params);
double to_double(const std::string& s, const std::string& what);
int to_int(const std::string& s, const std::string& what);

// Error as the reference raises it: a Messenger ERROR line, then an exception.
[[noreturn]] void fail(const MessengerPtr& msg, const std::string& error_line, const std::string& what);

// ---------------------------------------------------------------------------
// System: box, particles (input file) and the device context.
// ---------------------------------------------------------------------------
struct Box {
  double Lx = 10, Ly = 10, Lz = 10;
  double xlo() const { return -0.5 * Lx; } double xhi() const { return 0.5 * Lx; }
  double ylo() const { return -0.5 * Ly; } double yhi() const { return 0.5 * Ly; }
  double zlo() const { return -0.5 * Lz; } double zhi() const { return 0.5 * Lz; }
};

class System {
 public:
  // System(input, msg, box) (src/system/system.cpp:64-452): reads the particle
  // file ('keys:' header or the fixed column order) and uploads it.
  System(const std::string& input, MessengerPtr msg, const Box& box, bool periodic);
  ~System();
  samos_ctx* ctx() const { return ctx_; }
  int size() const { return n_; }
  int ntypes() const { return ntypes_; }
  const Box& box() const { return box_; }
  bool periodic() const { return periodic_; }
  void check(int rc) const;  // non-zero status -> ERROR + runtime_error
  double integrator_step() const { return dt_; }
  void set_integrator_step(double dt);
  int time_step = 0;  // cumulative across runs (System::get_step)
  std::vector<double> column(int field) const;
  std::vector<int> icolumn(int field) const;
  // System::make_group (src/system/system.cpp:462-570): criteria type,
  // not_type, r_eq, r_le, r_ge, x_le, x_ge, y_le, y_ge, z_le, z_ge on the
  // current state, all of the given ones; the members go to the device.
  void make_group(const std::string& name, pairs_type& params);
  bool has_group(const std::string& name) const { return groups_.count(name) != 0; }
  // Group::get_particles: member ids in the order they were added.
  const std::vector<int>& group(const std::string& name) const { return groups_.at(name); }

 private:
  std::map<std::string, std::vector<int>> groups_;
  samos_ctx* ctx_ = nullptr;
  MessengerPtr msg_;
  Box box_;
  bool periodic_ = true;
  int n_ = 0, ntypes_ = 1;
  double dt_ = 0.0;
};
using SystemPtr = std::shared_ptr<System>;

// ---------------------------------------------------------------------------
// Value (src/system/value.hpp:53-170): constant or linear schedules.
// ---------------------------------------------------------------------------
struct Value {
  int kind = SAMOS_VALUE_CONSTANT;
  double min_val = 0.0, max_val = 1.0;
  int steps = 1000;
};
using ValuePtr = std::shared_ptr<Value>;
ValuePtr make_value(const std::string& kind, MessengerPtr msg, pairs_type& params);

// ---------------------------------------------------------------------------
// Neighbour list (src/system/neighbour_list.hpp:124-272) - concrete.
// ---------------------------------------------------------------------------
class NeighbourList {
 public:
  NeighbourList(SystemPtr sys, MessengerPtr msg, double cutoff, double pad, pairs_type& params);
  double get_cutoff() const { return cut_; }
  void build();

 private:
  SystemPtr sys_;
  double cut_, pad_;
};
using NeighbourListPtr = std::shared_ptr<NeighbourList>;

// ---------------------------------------------------------------------------
// Pair potentials (src/potentials/pair/pair_potential.hpp:52-158).
// ---------------------------------------------------------------------------
class PairPotential {
 public:
  PairPotential(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, ValuePtr val, pairs_type& param);
  virtual ~PairPotential() = default;
  virtual void set_pair_parameters(pairs_type& pair_param) = 0;
  virtual bool need_nlist() { return true; }
  virtual void compute(double dt) = 0;        // adds this potential's forces on the device
  double get_potential_energy() const { return energy_; }
  std::string params_ok(pairs_type& params) const;

 protected:
  SystemPtr sys_;
  MessengerPtr msg_;
  NeighbourListPtr nlist_;
  ValuePtr val_;
  std::vector<std::string> known_;
  double energy_ = 0.0;
  int ntypes_ = 1;
};
using PairPotentialPtr = std::shared_ptr<PairPotential>;

class PairSoftPotential : public PairPotential {
 public:
  PairSoftPotential(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, ValuePtr val, pairs_type& param);
  void set_pair_parameters(pairs_type& pair_param) override;
  void compute(double dt) override;

 private:
  double k_ = 1.0, a_ = 2.0;
};

class PairLJPotential : public PairPotential {
 public:
  PairLJPotential(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, ValuePtr val, pairs_type& param);
  void set_pair_parameters(pairs_type& pair_param) override;
  void compute(double dt) override;

 private:
  double eps_ = 1.0, sigma_ = 1.0, rcut_ = 3.0;
};

// morse, gaussian, yukawa, soft_attractive: global parameters only (the
// per-type pair_param form is rejected).
class PairExtraPotential : public PairPotential {
 public:
  PairExtraPotential(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, ValuePtr val, pairs_type& param,
                     const std::string& name);
  void set_pair_parameters(pairs_type& pair_param) override;
  void compute(double dt) override;

 private:
  std::string name_;
};

// ---------------------------------------------------------------------------
// Pair aligners (src/aligner/pair/pair_aligner.hpp:53-89).
// ---------------------------------------------------------------------------
class PairAlign {
 public:
  PairAlign(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, pairs_type& param)
      : sys_(sys), msg_(msg), nlist_(nlist) {}
  virtual ~PairAlign() = default;
  virtual void set_pair_parameters(pairs_type& pair_param) = 0;
  virtual bool need_nlist() { return true; }
  virtual void compute() = 0;
  double get_potential_energy() const { return energy_; }

 protected:
  SystemPtr sys_;
  MessengerPtr msg_;
  NeighbourListPtr nlist_;
  double energy_ = 0.0;
  double J_ = 1.0, rcut_ = 0.0;
};
using PairAlignPtr = std::shared_ptr<PairAlign>;

class PairPolarAlign : public PairAlign {
 public:
  PairPolarAlign(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, pairs_type& param);
  void set_pair_parameters(pairs_type& pair_param) override;
  void compute() override;
};

class PairNematicAlign : public PairAlign {
 public:
  PairNematicAlign(SystemPtr sys, MessengerPtr msg, NeighbourListPtr nlist, pairs_type& param);
  void set_pair_parameters(pairs_type& pair_param) override;
  void compute() override;
};

// ---------------------------------------------------------------------------
// Constraints (src/constraints/constraint.hpp:47-161).
// ---------------------------------------------------------------------------
class Constraint {
 public:
  Constraint(SystemPtr sys, MessengerPtr msg, pairs_type& param);
  virtual ~Constraint() = default;
  virtual int kind() const = 0;
  virtual double param() const { return 0.0; }
  virtual int flag() const { return 0; }
  // Registers the constraint with the device context and enforces it once.
  virtual int install(samos_ctx* ctx) { return samos_constraint(ctx, kind(), param(), flag()); }

 protected:
  SystemPtr sys_;
  MessengerPtr msg_;
  int max_iter_ = 100;   // maxiter (constraint.hpp:62-73)
  double tol_ = 1e-6;    // tol (constraint.hpp:74-84)
  double rescale_ = 1.0; // rescale, rescale_steps, rescale_freq (constraint.hpp:85-108)
  int rescale_steps_ = 1000, rescale_freq_ = 10;
};
using ConstraintPtr = std::shared_ptr<Constraint>;

class ConstraintNone : public Constraint {
 public:
  ConstraintNone(SystemPtr sys, MessengerPtr msg, pairs_type& param);
  int kind() const override { return SAMOS_CONSTRAINT_NONE; }
};

class ConstraintPlane : public Constraint {
 public:
  ConstraintPlane(SystemPtr sys, MessengerPtr msg, pairs_type& param);
  int kind() const override { return SAMOS_CONSTRAINT_PLANE; }
  double param() const override { return zpos_; }
  int flag() const override { return unlimited_ ? 1 : 0; }

 private:
  double zpos_ = 0.0;
  bool unlimited_ = false;
};

class ConstraintSphere : public Constraint {
 public:
  ConstraintSphere(SystemPtr sys, MessengerPtr msg, pairs_type& param);
  int kind() const override { return SAMOS_CONSTRAINT_SPHERE; }
  double param() const override { return r_; }
  // ConstraintSphere::rescale (constraint_sphere.cpp:87-104) runs on the device
  int install(samos_ctx* ctx) override {
    int rc = samos_constraint(ctx, kind(), param(), flag());
    if (rc == 0 && rescale_ != 1.0) rc = samos_constraint_rescale(ctx, rescale_, rescale_steps_, rescale_freq_);
    return rc;
  }

 private:
  double r_ = 1.0;
};

// Cylinder {r}, torus {c, a} and ellipsoid {a, b, c}
// (src/constraints/constraint_cylinder.hpp, constraint_torus.hpp, constraint_ellipsoid.hpp).
class ConstraintSurface : public Constraint {
 public:
  ConstraintSurface(SystemPtr sys, MessengerPtr msg, pairs_type& param, const std::string& name);
  int kind() const override { return kind_; }
  int install(samos_ctx* ctx) override {
    return samos_constraint_surface(ctx, kind_, q_.data(), (int)q_.size(), max_iter_, tol_);
  }

 private:
  int kind_;
  std::vector<double> q_;
};

// ---------------------------------------------------------------------------
// Integrators (src/integrators/integrator.hpp:54-138).  Constructing one
// registers it with the device context, which then runs it in the run loop.
// ---------------------------------------------------------------------------
class Integrator {
 public:
  Integrator(SystemPtr sys, MessengerPtr msg, ValuePtr temp, pairs_type& param);
  virtual ~Integrator() = default;
  std::string params_ok(pairs_type& params) const;

 protected:
  void check_params(pairs_type& param, const std::string& what);
  SystemPtr sys_;
  MessengerPtr msg_;
  ValuePtr temp_;
  double dt_ = 0.0;
  std::vector<std::string> known_;
};
using IntegratorPtr = std::shared_ptr<Integrator>;

class IntegratorBrownian : public Integrator {
 public:
  IntegratorBrownian(SystemPtr sys, MessengerPtr msg, ValuePtr temp, pairs_type& param, int which);
};
class IntegratorLangevin : public Integrator {
 public:
  IntegratorLangevin(SystemPtr sys, MessengerPtr msg, ValuePtr temp, pairs_type& param);
};
class IntegratorFIRE : public Integrator {
 public:
  IntegratorFIRE(SystemPtr sys, MessengerPtr msg, ValuePtr temp, pairs_type& param);
};
class IntegratorNVE : public Integrator {
 public:
  IntegratorNVE(SystemPtr sys, MessengerPtr msg, ValuePtr temp, pairs_type& param);
};

// ---------------------------------------------------------------------------
// Registries (src/utils/factory_types.hpp:42-63, src/utils/register_*.cpp).
// ---------------------------------------------------------------------------
using PairPotentialFactory =
    std::function<PairPotentialPtr(SystemPtr, MessengerPtr, NeighbourListPtr, ValuePtr, pairs_type&)>;
using PairAlignFactory = std::function<PairAlignPtr(SystemPtr, MessengerPtr, NeighbourListPtr, pairs_type&)>;
using ConstraintFactory = std::function<ConstraintPtr(SystemPtr, MessengerPtr, pairs_type&)>;
using IntegratorFactory = std::function<IntegratorPtr(SystemPtr, MessengerPtr, ValuePtr, pairs_type&)>;

std::map<std::string, PairPotentialFactory>& pair_potentials();
std::map<std::string, PairAlignFactory>& pair_aligners();
std::map<std::string, ConstraintFactory>& constraints();
std::map<std::string, IntegratorFactory>& integrators();

// ---------------------------------------------------------------------------
// Observers (src/log/logger.cpp:40-137, src/dump/dump.cpp:351-690).
// ---------------------------------------------------------------------------
class Logger {
 public:
  Logger(SystemPtr sys, MessengerPtr msg, const std::string& file, pairs_type& params);
  bool due(int step) const { return step % freq_ == 0; }
  void log(int step);

 private:
  SystemPtr sys_;
  MessengerPtr msg_;
  std::ofstream out_;
  int freq_ = 100;
  std::vector<std::string> quantities_;
};
using LoggerPtr = std::shared_ptr<Logger>;

class Dump {
 public:
  Dump(SystemPtr sys, MessengerPtr msg, const std::string& name, pairs_type& params);
  bool due(int step) const { return step >= start_ && step % freq_ == 0; }
  void dump(int step);

 private:
  void write(std::ostream& out);
  void write_dcd_header();
  void write_dcd_frame();
  SystemPtr sys_;
  MessengerPtr msg_;
  std::string name_, type_, ext_, dir_ = ".", group_ = "all";
  pairs_type params_;
  int start_ = 0, freq_ = 1;
  double scale_ = 1.0;
  bool multi_ = false, header_ = false, keys_ = false;
  std::ofstream single_;
};
using DumpPtr = std::shared_ptr<Dump>;

}  // namespace samos_host
