// pmrun.cpp -- run a ProtoMol .conf file on the GPU.
//
//   pmrun <file.conf> [--keyword value ...] [--precision double|mixed]
//
// The sequence of applications/protomol-app/protomol.cpp:196-690: read the
// configuration (command-line "--keyword value" pairs override it), the
// positions, velocities (or random velocities at the temperature), PSF and
// PAR; build the topology; remove linear / angular momentum at start-up
// when asked; build the level-0 integrator and its forces; then alternate
// outputs and integration steps, with a final output at the last step.
// The integration runs in the device library (pm_md_run); outputs are
// allEnergiesFile, DCDFile, finXYZPosFile, finXYZVelFile and finPDBPosFile.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>

#include "pmf.h"
#include "pmgpu.h"

namespace {

const double kBoltzmann = 0.001987191;   // Constant::BOLTZMANN

int die(const std::string& msg) {
  std::fprintf(stderr, "pmrun: %s\n", msg.c_str());
  return 1;
}

struct Device {
  pm_context* c = nullptr;
  ~Device() {
    if (c) pm_destroy(c);
  }
  bool ok(int rc, std::string& err) const {
    if (rc == PM_OK) return true;
    err = pm_last_error(c);
    return false;
  }
};

bool vec3(const pmf::Config& cfg, const char* key, double out[3]) {
  if (!cfg.has(key)) return false;
  return std::sscanf(cfg.get(key).c_str(), "%lf %lf %lf", &out[0], &out[1], &out[2]) == 3;
}

// Replica exchange as applications/rem-app/rem.cpp:604-720 runs it: one
// replica per temperature of the REMTemperatureFile (one per line, in
// ladder order), NumSwitches rounds of numsteps steps, an exchange attempt
// after every round from EqSwitches on (neighbours of the round's parity,
// pm_batch_exchange); an accepted switch is credited to the replica of
// the higher temperature, and REMExchangeRatesFile holds each replica's
// credits divided by the rounds that attempted switches (the first line is
// always 0).  The two replicas of an accepted switch get new velocities at
// their new temperatures with linear and angular momentum removed (doSwitch,
// rem.cpp:906-936, which redraws with redrawVelocities = 1); the reference
// seeds that draw from the clock (getTimerSeed), here the seed is
// seed + 7919 * (round + 1) + replica, so runs repeat.  Every attempt prints
// one "REM attempt" line: the round, then per ladder slot the potential
// energy of the replica there before the attempt, then the accepted flag of
// each tested pair (k, k+1) by k.  All replicas run on this GPU as one batch.
template <class MakeCtx>
int run_rem(const pmf::Config& cfg, const pmf::System& sys, const std::vector<double>& x, const std::vector<double>& v,
            long lin, long ang, int integrator, MakeCtx& make_ctx) {
  std::string err;
  if (integrator != PM_INTEGRATOR_LANGEVIN) return die("replica exchange runs LangevinLeapfrog replicas");
  FILE* tf = std::fopen(cfg.path("remtemperaturefile").c_str(), "r");
  if (!tf) return die("cannot open REMTemperatureFile " + cfg.get("remtemperaturefile"));
  std::vector<double> temps;
  double t;
  while (std::fscanf(tf, "%lf", &t) == 1) temps.push_back(t);
  std::fclose(tf);
  if (temps.size() < 2) return die("REMTemperatureFile needs at least two temperatures");
  const int nsw = (int)cfg.integer("numswitches", -1);
  const int eq = (int)cfg.integer("eqswitches", 0);
  const int len = (int)cfg.integer("numsteps", 0);
  if (nsw < 1 || len < 1) return die("replica exchange needs NumSwitches >= 1 and numsteps >= 1");
  const int R = (int)temps.size();
  std::vector<Device> devs(R);
  std::vector<pm_context*> ctxs(R);
  std::vector<int> handles;
  for (int r = 0; r < R; r++) {
    // every replica starts from the .conf's positions with velocities at its own temperature
    std::vector<double> vr = v;
    if (!cfg.has("velfile")) {
      vr = pmf::random_velocity(sys, temps[r], (unsigned)cfg.integer("seed", 1234));
      if (lin >= 0 || ang >= 0) pmf::remove_momentum(sys, x, vr, lin >= 0, ang >= 0);
    }
    if (!make_ctx(devs[r], temps[r], x, vr, handles, err)) return die(err);
    ctxs[r] = devs[r].c;
  }
  pm_batch* batch = nullptr;
  if (pm_batch_create(ctxs.data(), R, &batch) != PM_OK) return die("pm_batch_create failed");
  std::unique_ptr<pm_batch, void (*)(pm_batch*)> hold(batch, pm_batch_destroy);
  if (pm_batch_set_exchange_velocities(batch, PM_EXCHANGE_KEEP) != PM_OK) return die("pm_batch_set_exchange_velocities");
  std::vector<double> credit(R, 0.0);
  std::vector<int> acc(R, 0), ladder(R);
  std::vector<double> slot_t(R), xr(3 * (size_t)sys.n), e(PM_ENERGY_COUNT);
  const unsigned long long seed = (unsigned long long)cfg.integer("seed", 1234);
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < nsw; i++) {
    if (pm_batch_run(batch, len) != PM_OK) return die(std::string("pm_batch_run: ") + pm_batch_last_error(batch));
    if (i < eq) continue;
    // the potential energies the attempt compares, by ladder slot
    if (pm_batch_ladder(batch, ladder.data(), slot_t.data()) != PM_OK) return die("pm_batch_ladder");
    std::printf("REM attempt %d", i);
    for (int k = 0; k < R; k++) {
      Device tmp;   // not owned: only for the error text
      tmp.c = ctxs[ladder[k]];
      const bool ok = tmp.ok(pm_md_get_state(ctxs[ladder[k]], nullptr, nullptr, e.data()), err);
      tmp.c = nullptr;
      if (!ok) return die(err);
      double ep = 0;
      for (int u = 0; u < 7; u++) ep += e[u];   // ScalarStructure::potentialEnergy
      std::printf(" %.17g", ep);
    }
    std::fill(acc.begin(), acc.end(), 0);
    if (pm_batch_exchange(batch, i, seed, acc.data(), nullptr) != PM_OK)
      return die(std::string("pm_batch_exchange: ") + pm_batch_last_error(batch));
    for (int k = (i & 1); k + 1 < R; k += 2) std::printf(" %d", acc[k]);
    std::printf("\n");
    if (pm_batch_ladder(batch, ladder.data(), slot_t.data()) != PM_OK) return die("pm_batch_ladder");
    for (int k = (i & 1); k + 1 < R; k += 2) {
      if (!acc[k]) continue;
      credit[k + 1] += 1.0;
      // doSwitch: randomVelocity at the slot's temperature, then
      // removeLinearMomentum and removeAngularMomentum
      for (int u = k; u <= k + 1; u++) {
        const int r = ladder[u];
        Device tmp;
        tmp.c = ctxs[r];
        bool ok = tmp.ok(pm_md_get_state(ctxs[r], xr.data(), nullptr, nullptr), err);
        if (ok) {
          std::vector<double> vr = pmf::random_velocity(sys, slot_t[u], (unsigned)(seed + 7919ull * (i + 1) + r));
          pmf::remove_momentum(sys, xr, vr, true, true);
          ok = tmp.ok(pm_md_set_state(ctxs[r], nullptr, vr.data()), err);
        }
        tmp.c = nullptr;
        if (!ok) return die(err);
      }
    }
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (cfg.has("remexchangeratesfile")) {
    FILE* f = std::fopen(cfg.path("remexchangeratesfile").c_str(), "w");
    if (!f) return die("cannot open " + cfg.path("remexchangeratesfile"));
    for (int r = 0; r < R; r++) std::fprintf(f, "%g\n", credit[r] / (nsw - eq));
    std::fclose(f);
  }
  std::printf("Exchange counts:");
  for (int r = 0; r < R; r++) std::printf(" %g", credit[r]);
  std::printf("\nREM: %d replicas x %d switches x %d steps in %.3f s (%.0f replica-steps/s)\n", R, nsw, len, secs,
              secs > 0 ? (double)R * nsw * len / secs : 0.0);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: pmrun <file.conf> [--keyword value ...] [--precision double|mixed]\n");
    return 2;
  }
  std::string err;
  pmf::Config cfg;
  if (!pmf::read_conf(argv[1], cfg, err)) return die(err);
  if (!pmf::apply_overrides(cfg, argc - 2, argv + 2, err)) return die(err);
  for (const auto& w : cfg.ignored) std::printf("Ignoring: %s\n", w.c_str());

  // ---------------------------------------------------------- input files
  if (!cfg.has("posfile") || !cfg.has("psffile") || !cfg.has("parfile"))
    return die("posfile, psffile and parfile are required");
  pmf::Coords pos, vel;
  if (!pmf::read_coords(cfg.path("posfile"), pos, err)) return die(err);
  std::printf("Using %s posfile '%s' (%zu).\n", pos.format.c_str(), cfg.get("posfile").c_str(), pos.xyz.size() / 3);
  pmf::Psf psf;
  pmf::Par par;
  if (!pmf::read_psf(cfg.path("psffile"), psf, err)) return die(err);
  if (!pmf::read_par(cfg.path("parfile"), par, pmf::PAR_AUTO, err)) return die(err);
  const bool mult = cfg.has("dihedralmultpsf") ? cfg.boolean("dihedralmultpsf", false) : par.dialect != pmf::PAR_CHARMM28;
  int exclude = pmf::EXCL_SCALED14;
  if (cfg.has("exclude") && !pmf::parse_exclude(cfg.get("exclude"), exclude))
    return die("exclude must be none, 1-2, 1-3, 1-4 or scaled1-4");
  pmf::System sys;
  if (!pmf::build_system(psf, par, exclude, mult, sys, err)) return die(err);
  const int n = sys.n;
  std::printf("Using PSF file '%s' (%d), PAR file '%s' (%s charmm force field).\n", cfg.get("psffile").c_str(), n,
              cfg.get("parfile").c_str(), par.dialect == pmf::PAR_CHARMM28 ? "new" : "old");
  if ((int)pos.xyz.size() != 3 * n) return die("positions and PSF have different numbers of atoms");
  std::vector<double> x = pos.xyz, v;
  if (cfg.has("velfile")) {
    if (!pmf::read_coords(cfg.path("velfile"), vel, err)) return die(err);
    if ((int)vel.xyz.size() != 3 * n) return die("velocities and PSF have different numbers of atoms");
    v = vel.xyz;
  } else {
    if (!cfg.has("temperature")) return die("velfile or temperature is required");
    v = pmf::random_velocity(sys, cfg.real("temperature", 300.0), (unsigned)cfg.integer("seed", 1234));
  }
  const long lin = cfg.has("removelinearmomentum") ? cfg.integer("removelinearmomentum", -1) : cfg.integer("commotion", -1);
  const long ang = cfg.has("removeangularmomentum") ? cfg.integer("removeangularmomentum", -1)
                                                     : cfg.integer("angularmomentum", -1);
  if (lin > 0 || ang > 0)
    return die("removeLinearMomentum / removeAngularMomentum > 0 (removal every n steps) is not implemented; "
               "0 removes at start-up, -1 never");
  if (lin >= 0 || ang >= 0) pmf::remove_momentum(sys, x, v, lin >= 0, ang >= 0);

  // ------------------------------------------------------------- boundary
  const std::string bc = cfg.get("boundaryconditions", "vacuum");
  double e1[3], e2[3], e3[3], org[3] = {0, 0, 0};
  if (pmf::upper(bc) != "PERIODIC") return die("boundaryConditions " + bc + ": the GPU path runs periodic boxes");
  if (!vec3(cfg, "cellbasisvector1", e1) || !vec3(cfg, "cellbasisvector2", e2) || !vec3(cfg, "cellbasisvector3", e3))
    return die("periodic boundary conditions need cellBasisVector1, 2 and 3");
  vec3(cfg, "cellorigin", org);
  const double cellsize = cfg.real("cellsize", 6.5);
  const double volume = std::fabs(e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                                  e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]));

  // ----------------------------------------------------------- integrator
  if (cfg.levels.size() != 1 || cfg.levels[0].level != 0)
    return die("one integrator level (level 0) is supported on the GPU path");
  const pmf::LevelSpec& L = cfg.levels[0];
  auto lparam = [&](const char* k, double def) {
    for (const auto& p : L.params)
      if (p.first == k) {
        double d;
        if (pmf::to_real(p.second, d)) return d;
      }
    return def;
  };
  int integrator;
  const std::string iname = pmf::upper(L.integrator);
  if (iname == "LEAPFROG") integrator = PM_INTEGRATOR_LEAPFROG;
  else if (iname == "LANGEVINLEAPFROG") integrator = PM_INTEGRATOR_LANGEVIN;
  else return die("integrator " + L.integrator + " is not implemented on the GPU (Leapfrog, LangevinLeapfrog)");
  const double dt = lparam("timestep", -1.0);
  if (!(dt > 0)) return die("the integrator needs a positive timestep");

  // --------------------------------------------------------------- device
  const std::string prec = cfg.get("precision", "double");
  const bool shake = cfg.boolean("shake", false) && cfg.real("shakeepsilon", 1e-5) > 0 && cfg.integer("shakemaxiter", 30) > 0;
  const bool rattle =
      cfg.boolean("rattle", false) && cfg.real("rattleepsilon", 1e-5) > 0 && cfg.integer("rattlemaxiter", 30) > 0;
  int nconstraints = 0;
  // one context with the topology, forces and integrator of the .conf at a
  // bath temperature (the level's own for plain runs, the ladder's for REM)
  auto make_ctx = [&](Device& dev, double temperature, const std::vector<double>& x0, const std::vector<double>& v0,
                      std::vector<int>& handles, std::string& err) -> bool {
    if (pm_create(0, &dev.c) != PM_OK) {
      err = "no HIP device";
      return false;
    }
    if (!dev.ok(pm_set_precision(dev.c, prec == "mixed" ? PM_PRECISION_MIXED : PM_PRECISION_DOUBLE), err)) return false;
    const int nt = (int)sys.type_name.size();
    if (!dev.ok(pm_set_topology(dev.c, n, sys.q.data(), sys.type.data(), nt, sys.ljA.data(), sys.ljB.data(),
                                sys.ljA14.data(), sys.ljB14.data(), sys.molecule.data(), sys.mass.data(),
                                (int)sys.ex_a.size(), sys.ex_a.data(), sys.ex_b.data(), sys.ex_cls.data(), 1.0),
                err))
      return false;
    if (!sys.bond_a.empty() || !sys.angle_a.empty())
      if (!dev.ok(pm_set_bonded(dev.c, (int)sys.bond_a.size(), sys.bond_a.data(), sys.bond_b.data(),
                                sys.bond_k.data(), sys.bond_r0.data(), (int)sys.angle_a.size(), sys.angle_a.data(),
                                sys.angle_b.data(), sys.angle_c.data(), sys.angle_k.data(), sys.angle_t0.data(),
                                sys.angle_kub.data(), sys.angle_rub.data()),
                  err))
        return false;
    for (int kind = 0; kind < 2; kind++) {
      const auto& list = kind ? sys.impropers : sys.dihedrals;
      if (list.empty()) continue;
      std::vector<int> at, row{0}, per;
      std::vector<double> k, ph;
      for (const auto& t : list) {
        at.insert(at.end(), t.a, t.a + 4);
        k.insert(k.end(), t.k.begin(), t.k.end());
        per.insert(per.end(), t.n.begin(), t.n.end());
        ph.insert(ph.end(), t.phase.begin(), t.phase.end());
        row.push_back((int)k.size());
      }
      if (!dev.ok(pm_set_torsions(dev.c, kind, (int)list.size(), at.data(), row.data(), k.data(), per.data(),
                                  ph.data()),
                  err))
        return false;
    }
    if (!dev.ok(pm_set_boundary(dev.c, e1, e2, e3, org, cellsize), err)) return false;
    handles.clear();
    for (const auto& fs : L.forces) {
      std::string id;
      std::vector<double> vals;
      if (!pmf::force_values(fs, id, vals, err)) return false;
      int h;
      if (!dev.ok(pm_force_create(dev.c, id.c_str(), vals.data(), (int)vals.size(), &h), err)) return false;
      handles.push_back(h);
    }
    if (!dev.ok(pm_md_setup(dev.c, handles.data(), (int)handles.size(), integrator, dt, temperature,
                            lparam("gamma", 0.0), (unsigned long long)lparam("seed", 1234)),
                err))
      return false;
    if (shake || rattle) {
      if (!dev.ok(pm_md_set_constraints(dev.c, shake ? cfg.real("shakeepsilon", 1e-5) : 0.0,
                                        shake ? (int)cfg.integer("shakemaxiter", 30) : 0,
                                        rattle ? cfg.real("rattleepsilon", 1e-5) : 0.0,
                                        rattle ? (int)cfg.integer("rattlemaxiter", 30) : 0),
                  err))
        return false;
      pm_md_constraint_count(dev.c, &nconstraints);
    }
    return dev.ok(pm_md_set_state(dev.c, x0.data(), v0.data()), err);
  };
  for (const auto& fs : L.forces) {
    std::string id;
    std::vector<double> vals;
    if (!pmf::force_values(fs, id, vals, err)) return die(err);
    std::printf("Force %s", id.c_str());
    for (double d : vals) std::printf(" %g", d);
    std::printf("\n");
  }
  if (cfg.has("remtemperaturefile")) return run_rem(cfg, sys, x, v, lin, ang, integrator, make_ctx);
  Device dev;
  std::vector<int> handles;
  if (!make_ctx(dev, lparam("temperature", 300.0), x, v, handles, err)) return die(err);
  const int dof = sys.dof - nconstraints;

  // -------------------------------------------------------------- outputs
  const int first = (int)cfg.integer("firststep", 0);
  const int last = first + (int)cfg.integer("numsteps", 0);
  const int freq = std::max(1, (int)cfg.integer("outputfreq", 1));
  FILE* ef = nullptr;
  int efreq = freq, enext = first;
  if (cfg.has("allenergiesfile")) {
    ef = std::fopen(cfg.path("allenergiesfile").c_str(), "w");
    if (!ef) return die("cannot open " + cfg.path("allenergiesfile"));
    efreq = std::max(1, (int)cfg.integer("allenergiesfileoutputfreq", freq));
    FILE* hf = std::fopen((cfg.path("allenergiesfile") + ".header").c_str(), "w");
    if (hf) {
      const char* cols[] = {"Time(fs)", "E_potential", "E_kinetic", "E_total", "Temperature", "E_bond", "E_angle",
                            "E_dihedral", "E_improper", "E_VdW", "E_coulomb", "E_other", "Volume(A^3)"};
      for (int c = 0; c < 13; c++) std::fprintf(hf, c ? " %14s" : "%14s", cols[c]);
      std::fprintf(hf, "\n");
      std::fclose(hf);
    }
  }
  pmf::DcdWriter dcd;
  int dfreq = freq, dnext = first;
  const bool do_dcd = cfg.has("dcdfile");
  if (do_dcd) {
    dfreq = std::max(1, (int)cfg.integer("dcdfileoutputfreq", freq));
    if (!dcd.open(cfg.path("dcdfile"), n, 1, dfreq, 1.0, "", err)) return die(err);
  }
  std::vector<double> e(PM_ENERGY_COUNT, 0.0);
  // energies at the start positions (the integrator's initial force evaluation)
  {
    std::vector<double> f(3 * (size_t)n);
    if (!dev.ok(pm_evaluate(dev.c, handles.data(), (int)handles.size(), x.data(), f.data(), e.data(), 0), err))
      return die(err);
  }
  auto emit = [&](int step) {
    double ke = 0;
    for (int i = 0; i < n; i++)
      ke += 0.5 * sys.mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    double pe = 0;
    for (int k = 0; k < 7; k++) pe += e[k];
    pmf::EnergyRow r;
    r.time_fs = step * dt;
    r.potential = pe;
    r.kinetic = ke;
    r.total = pe + ke;
    r.temperature = 2.0 * ke / (kBoltzmann * dof);
    r.coulomb = e[0]; r.lj = e[1]; r.bond = e[2]; r.angle = e[3]; r.dihedral = e[4]; r.improper = e[5]; r.other = e[6];
    r.volume = volume;
    if (ef) {
      std::fputs(pmf::energies_line(r).c_str(), ef);
      std::fflush(ef);
    }
    return r;
  };
  auto run_outputs = [&](int step, bool final_) {
    bool any = false;
    if (ef && (final_ || step >= enext)) {
      if (!final_) enext += std::max((step - enext) / efreq, 1) * efreq;
      any = true;
    }
    if (any) emit(step);
    if (do_dcd && (final_ || step >= dnext)) {
      if (!final_) dnext += std::max((step - dnext) / dfreq, 1) * dfreq;
      if (!dcd.frame(x, err)) std::fprintf(stderr, "pmrun: %s\n", err.c_str());
    }
  };
  const auto t0 = std::chrono::steady_clock::now();
  int step = first;
  while (step < last) {
    run_outputs(step, false);
    int next = std::min(ef ? enext : last, do_dcd ? dnext : last);
    int inc = std::min(last, std::max(next, step + 1)) - step;
    if (!dev.ok(pm_md_run(dev.c, inc), err)) return die(err);
    step += inc;
    if (!dev.ok(pm_md_get_state(dev.c, x.data(), v.data(), e.data()), err)) return die(err);
    if (step % 1000 < inc) std::printf("step %d\n", step), std::fflush(stdout);
  }
  run_outputs(last, true);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (ef) std::fclose(ef);
  dcd.close();
  const std::string comment = "Time : " + std::to_string(last * dt) + ", step : " + std::to_string(last) + ".";
  std::vector<std::string> names(n);
  for (int i = 0; i < n; i++) names[i] = sys.type_name[sys.type[i]];
  if (cfg.has("finxyzposfile") && !pmf::write_xyz(cfg.path("finxyzposfile"), x, names, comment, err)) return die(err);
  if (cfg.has("finxyzvelfile") && !pmf::write_xyz(cfg.path("finxyzvelfile"), v, names, comment, err)) return die(err);
  if (cfg.has("finpdbposfile") && !pmf::write_pdb(cfg.path("finpdbposfile"), x, pos.pdb, comment, err)) return die(err);
  const double ns_day = secs > 0 ? (last - first) * dt * 1e-6 / (secs / 86400.0) : 0.0;
  std::printf("Run: %d steps of %g fs on %d atoms in %.3f s (%.3f ms/step, %.2f ns/day, %s)\n", last - first, dt, n,
              secs, secs * 1e3 / std::max(1, last - first), ns_day, prec.c_str());
  return 0;
}
