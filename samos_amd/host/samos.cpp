// `samos <conf>`: the configuration-file driver (src/samos.cpp:50-1388) for the
// particle pipeline.  Reads the conf line by line, builds the plugins through
// the registries, and runs `run N` on the device, stopping at every step where
// a log or dump is due so observers see the reference's loop-head state.
#include <chrono>
#include <cstdio>
#include <iostream>

#include "host.hpp"

using namespace samos_host;

namespace {

constexpr double kDefaultCutoff = 3.0;  // src/system/defaults.hpp:36-37
constexpr double kDefaultPadding = 0.5;

// Commands of the reference grammar (src/parser/parse_command.hpp:66-102) that
// belong to models outside this implementation.
const char* kOutOfScope[] = {"external_param", "external_align", "external_align_param", "population",
                             "zero_momentum", "bond", "bond_param", "angle", "angle_param", "read_bonds",
                             "read_angles", "read_cell_boundary", "pair_type_param", "disable_population"};

struct Driver {
  MessengerPtr msg = std::make_shared<Messenger>();
  Box box;
  bool periodic = true, have_box = false, have_timestep = false;
  SystemPtr sys;
  NeighbourListPtr nlist;
  std::map<std::string, PairPotentialPtr> pots;
  std::map<std::string, PairAlignPtr> aligns;
  ConstraintPtr constraint;
  std::map<std::string, IntegratorPtr> integs;
  std::vector<LoggerPtr> logs;
  std::vector<DumpPtr> dumps;
  long long total_steps = 0;
  double run_seconds = 0.0;

  void need_system(const std::string& what) {
    if (!sys)
      fail(msg, "System has not been defined. Please define system using \"input\" command before " + what + ".",
           "System not defined.");
  }

  void default_nlist() {
    if (nlist) return;
    msg->msg(Messenger::WARNING, "No neighbour list defined. Some pair potentials (e.g. Lennard-Jones) need it. "
                                 "We are making one with default cutoff and padding.");
    pairs_type none;
    nlist = std::make_shared<NeighbourList>(sys, msg, kDefaultCutoff, kDefaultPadding, none);
  }

  void command(const CommandLine& c, int line_no) {
    std::string attrib;
    pairs_type p;
    const std::string where = " at line " + std::to_string(line_no) + ".";
    auto parse = [&]() {
      if (!parse_attrib_params(c.rest, attrib, p))
        fail(msg, "Could not parse parameters of command " + c.command + where, "Error parsing command script.");
    };
    const std::string& cmd = c.command;
    if (cmd == "messages") {
      parse();
      msg->set_file(attrib);
    } else if (cmd == "config") {
      msg->msg(Messenger::WARNING, "config: the parameter record file is not written by this implementation.");
    } else if (cmd == "box") {
      parse();
      if (attrib == "periodic") periodic = true;
      else if (attrib == "fixed") periodic = false;
      else fail(msg, "Box type " + attrib + " is not known.", "Unknown box type.");
      if (p.count("lx")) box.Lx = to_double(p["lx"], "lx");
      if (p.count("ly")) box.Ly = to_double(p["ly"], "ly");
      if (p.count("lz")) box.Lz = to_double(p["lz"], "lz");
      have_box = true;
    } else if (cmd == "input") {
      parse();
      sys = std::make_shared<System>(attrib, msg, box, periodic);
    } else if (cmd == "nlist") {
      need_system("adding neighbour list");
      parse();
      const double rcut = p.count("rcut") ? to_double(p["rcut"], "rcut") : kDefaultCutoff;
      const double pad = p.count("pad") ? to_double(p["pad"], "pad") : kDefaultPadding;
      nlist = std::make_shared<NeighbourList>(sys, msg, rcut, pad, p);
      msg->msg(Messenger::INFO, "Created neighbour list with cutoff " + std::to_string(rcut) + " and padding distance " +
                                    std::to_string(pad) + ".");
    } else if (cmd == "constraint") {
      need_system("adding constraint");
      parse();
      if (constraint)
        fail(msg, "More than one constraint is not supported by this implementation.", "Unsupported constraint combination.");
      auto it = constraints().find(attrib);
      if (it == constraints().end())
        fail(msg, "Constraint " + attrib + " is not available (sphere, plane, none, cylinder, torus, ellipsoid)" + where,
             "Unknown constraint.");
      constraint = it->second(sys, msg, p);
      sys->check(constraint->install(sys->ctx()));
      msg->msg(Messenger::INFO, "Adding constraint of type " + attrib + ".");
    } else if (cmd == "pair_potential") {
      need_system("adding any pair potentials");
      parse();
      auto it = pair_potentials().find(attrib);
      if (it == pair_potentials().end())
        fail(msg, "Pair potential " + attrib + " is not available (soft, lj, morse, gaussian, yukawa, soft_attractive)" + where, "Unknown pair potential.");
      default_nlist();
      const std::string phase_in = p.count("phase_in") ? p["phase_in"] : "constant";
      ValuePtr v = make_value(phase_in.empty() ? "constant" : phase_in, msg, p);
      pots[attrib] = it->second(sys, msg, nlist, v, p);
      msg->msg(Messenger::INFO, "Added " + attrib + " to the list of pair potentials.");
    } else if (cmd == "pair_param") {
      parse();
      if (!pots.count(attrib))
        fail(msg, "No pair potentials have been defined. Please define them using \"pair_potential\" command before "
                  "modifying any pair parameters.", "No pair potentials defined.");
      pots[attrib]->set_pair_parameters(p);
    } else if (cmd == "pair_align") {
      need_system("adding any pair alignment");
      parse();
      auto it = pair_aligners().find(attrib);
      if (it == pair_aligners().end())
        fail(msg, "Pair aligner " + attrib + " is not available (polar, nematic)" + where, "Unknown pair aligner.");
      default_nlist();
      aligns[attrib] = it->second(sys, msg, nlist, p);
    } else if (cmd == "align_param") {
      parse();
      if (!aligns.count(attrib))
        fail(msg, "No pair alignment has been defined. Please define it using \"pair_align\" command before "
                  "modifying any alignment parameters.", "No pair alignment defined.");
      aligns[attrib]->set_pair_parameters(p);
    } else if (cmd == "external") {
      need_system("adding external potentials");
      parse();
      if (attrib != "self_propulsion")
        fail(msg, "External potential " + attrib + " is not available (self_propulsion)" + where, "Unknown external potential.");
      const double alpha = p.count("alpha") ? to_double(p["alpha"], "alpha") : 1.0;
      sys->check(samos_external_self_propulsion(sys->ctx(), alpha));
    } else if (cmd == "timestep") {
      need_system("setting the time step");
      sys->set_integrator_step(to_double(c.rest, "timestep"));
      have_timestep = true;
    } else if (cmd == "integrator") {
      need_system("adding the integrator");
      parse();
      if (!constraint)
        fail(msg, "Constraint has not been defined. Please define them using \"constraint\" command before adding "
                  "integrator.", "Constraint not defined.");
      if (!have_timestep && !p.count("dt"))
        fail(msg, "Simulation timestep has not been defined  Please define them using the \"timestep\" command before "
                  "adding integrator.", "Timestep not defined.");
      auto it = integrators().find(attrib);
      if (it == integrators().end())
        fail(msg, "Integrator " + attrib + " is not available (brownian, brownian_pos, brownian_align, langevin, fire)" +
                      where, "Unknown integrator.");
      const std::string group = p.count("group") ? p["group"] : "all";
      const std::string key = attrib + "_" + group;
      if (integs.count(key))
        fail(msg, "Integrator of type " + attrib + " (for group " + group + ") already exists. You need to disable "
                  "the old one first.", "Integrator already exists.");
      const std::string tc = p.count("temperature_control") ? p["temperature_control"] : "constant";
      integs[key] = it->second(sys, msg, make_value(tc, msg, p), p);
      msg->msg(Messenger::INFO, "Adding integrator of type " + attrib + " (for group " + group + ").");
    } else if (cmd == "disable") {
      parse();
      const std::string group = p.count("group") ? p["group"] : "all";
      if (!integs.count(attrib + "_" + group))
        fail(msg, "Integrator of type " + attrib + " (for group " + group + ") does not exist.", "Integrator does not exist.");
      sys->check(samos_integrator_disable(sys->ctx(), attrib.c_str(), group.c_str()));
      integs.erase(attrib + "_" + group);
    } else if (cmd == "log") {
      need_system("adding any logs");
      parse();
      logs.push_back(std::make_shared<Logger>(sys, msg, attrib, p));
    } else if (cmd == "dump") {
      need_system("adding any dumps");
      parse();
      dumps.push_back(std::make_shared<Dump>(sys, msg, attrib, p));
    } else if (cmd == "group") {
      if (!sys)
        fail(msg, "System has not been defined. Please define system using \"input\" command before adding any "
                  "particle groups.", "System not defined.");
      parse();
      sys->make_group(attrib, p);
      msg->msg(Messenger::INFO, "Adding group " + attrib + ".");
    } else if (cmd == "ntypes") {
      msg->msg(Messenger::INFO, "ntypes: per-type tables hold 16 types.");
    } else if (cmd == "run") {
      run(to_int(c.rest, "run"));
    } else {
      for (const char* o : kOutOfScope)
        if (cmd == o)
          fail(msg, "Command " + cmd + " belongs to a model outside this implementation" + where, "Unsupported command.");
      std::cerr << "Error parsing line : " << line_no << std::endl;
      std::cerr << "Unknown command : " << cmd << std::endl;
      throw std::runtime_error("Error parsing command script.");
    }
  }

  // The `run N` loop (src/samos.cpp:837-977): N+1 iterations; observers run at
  // the loop head of every iteration where one of them is due.
  void run(int steps) {
    if (!sys)
      fail(msg, "System has not been defined. Please define system using \"input\" command before adding any dumps.",
           "System not defined.");
    if (!constraint)
      fail(msg, "Constraint has not been defined. Please define it using \"constraint\" command before running "
                "simulation.", "Constraint not defined.");
    if (integs.empty())
      fail(msg, "Integrator has not been defined. Please define it using \"integrator\" command before running "
                "simulation.", "Integrator not defined.");
    msg->msg(Messenger::INFO, "Starting simulation run for " + std::to_string(steps) + " steps.");
    auto t0 = std::chrono::steady_clock::now();
    const int ts0 = sys->time_step;
    int done = 0;
    for (int t = 0; t <= steps; ++t) {
      const int ts = ts0 + t;
      bool due = false;
      for (auto& l : logs) due |= l->due(ts);
      for (auto& d : dumps) due |= d->due(ts);
      if (!due) continue;
      sys->check(samos_run_range(sys->ctx(), done, t, steps, 0));
      done = t;
      for (auto& d : dumps) d->dump(ts);
      for (auto& l : logs) l->log(ts);
    }
    sys->check(samos_run_range(sys->ctx(), done, steps + 1, steps, 0));
    sys->time_step = ts0 + steps + 1;
    long long info[8];
    sys->check(samos_nlist_info(sys->ctx(), info));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    total_steps += steps + 1;
    run_seconds += secs;
    msg->msg(Messenger::INFO, "Built neighbour list " + std::to_string(info[5]) + " times in total.");
    msg->msg(Messenger::INFO, "Run of " + std::to_string(steps + 1) + " steps took " + std::to_string(secs) + " s.");
  }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "Usage: " << argv[0] << " <configuration file>" << std::endl;
    return 1;
  }
  std::ifstream conf(argv[1]);
  if (!conf) {
    std::cerr << "Could not open configuration file " << argv[1] << std::endl;
    return 1;
  }
  Driver d;
  auto start = std::chrono::steady_clock::now();
  try {
    std::string line;
    int line_no = 1;
    while (std::getline(conf, line)) {
      CommandLine c;
      if (parse_command(line, c)) d.command(c, line_no);
      ++line_no;
    }
  } catch (const std::exception& e) {
    std::cerr << "samos: " << e.what() << std::endl;
    return 1;
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
  std::printf("Total run time : %.3f s\n", secs);
  if (d.run_seconds > 0)
    std::printf("Average %.1f time steps per second (%.4e particle-steps/s)\n", d.total_steps / d.run_seconds,
                d.sys ? d.total_steps * (double)d.sys->size() / d.run_seconds : 0.0);
  return 0;
}
