// pmf_run.cpp -- what the protomol application does between reading the
// files and running the integrator: force definitions turned into the
// device library's force ids and positional values, and the start-up
// removal of linear and angular momentum.
#include "pmf.h"

#include <cmath>
#include <cstdlib>
#include <set>
#include <sstream>

namespace pmf {

namespace {

std::string lower(const std::string& s) {
  std::string u(s);
  for (auto& ch : u) ch = (char)std::tolower((unsigned char)ch);
  return u;
}

struct Param {
  std::string key;    // "" for a value that follows the previous one
  bool has_default;
  double def;
};

// topologyutilities.cpp atomTypeToSymbolName; equalBeginNocase compares the
// shorter of the two words with the start of the other
bool begins_either(const std::string& a, const std::string& b) {
  const size_t n = std::min(a.size(), b.size());
  for (size_t i = 0; i < n; i++)
    if (std::toupper((unsigned char)a[i]) != std::toupper((unsigned char)b[i])) return false;
  return true;
}
std::string symbol_of(const std::string& type) {
  static const char* map[][2] = {{"SOD", "NA"}, {"POT", "K"}, {"CLA", "CL"}, {"CAL", "CA"}, {"CES", "CS"},
                                 {"FE", "FE"},  {"HE", "HE"}, {"NE", "NE"},   {"H", "H"},    {"C", "C"},
                                 {"O", "O"},    {"F", "F"},   {"N", "N"},     {"P", "P"},    {"S", "S"},
                                 {"MG", "MG"},  {"ZN", "ZN"}};
  for (auto& m : map)
    if (begins_either(type, m[0])) return m[1];
  return type;
}

}  // namespace

// ForceFactoryDetails::doMake (frontend/ForceFactoryDetails.cpp:48-190): the
// words of the definition that name the force (names, -algorithm,
// -switchingFunction, -interpolation and the PME part flags) give the id;
// every other word is a parameter word.  Each parameter of the force's
// list is looked up by its keyword from where the previous one ended,
// wrapping around once; a keyword-less parameter takes the word after the
// previous value; a parameter not found takes its default.
bool force_values(const ForceSpec& fs, std::string& id, std::vector<double>& values, std::string& err) {
  std::vector<std::string> words;
  {
    std::istringstream ss(fs.id);
    std::string t;
    while (ss >> t) words.push_back(t);
  }
  // "time" and "compare" wrap a force for timing / comparison only
  while (!words.empty() && (lower(words[0]) == "time" || lower(words[0]) == "compare")) words.erase(words.begin());
  std::vector<std::string> names, swfs;
  std::string algo, interp;
  bool real = false, recip = false, corr = false;
  std::set<std::string> idwords;
  size_t k = 0;
  while (k < words.size() && words[k][0] != '-') {
    names.push_back(words[k]);
    idwords.insert(lower(words[k]));
    k++;
  }
  for (size_t i = k; i < words.size(); i++) {
    const std::string w = lower(words[i]);
    if ((w == "-algorithm" || w == "-switchingfunction" || w == "-interpolation") && i + 1 < words.size()) {
      const std::string v = words[++i];
      if (w == "-algorithm") algo = v;
      else if (w == "-switchingfunction") swfs.push_back(v);
      else interp = v;
      idwords.insert(w);
      idwords.insert(lower(v));
    } else if (w == "-real" || w == "-reciprocal" || w == "-correction") {
      if (w == "-real") real = true;
      if (w == "-reciprocal") recip = true;
      if (w == "-correction") corr = true;
      idwords.insert(w);
    }
  }
  if (names.empty()) {
    err = "force definition '" + fs.id + "' names no force";
    return false;
  }
  std::vector<std::string> pw;
  for (const auto& w : words)
    if (!idwords.count(lower(w))) pw.push_back(w);

  std::vector<Param> params;
  std::string n0 = lower(names[0]);
  values.clear();
  if (names.size() == 1 && (n0 == "bond" || n0 == "angle" || n0 == "dihedral" || n0 == "improper")) {
    id = names[0];
    return true;
  }
  if (lower(algo) == "pmewald") {
    id = names[0] + " -algorithm PMEwald";
    if (!real && !recip && !corr) real = recip = corr = true;
    if (real) id += " -real";
    if (recip) id += " -reciprocal";
    if (corr) id += " -correction";
    id += " -interpolation " + (interp.empty() ? std::string("BSpline") : interp);
    params = {{"-gridsize", false, 0}, {"", false, 0}, {"", false, 0}, {"-cutoff", false, 0},
              {"-order", true, 4}, {"-accuracy", true, 1e-6}, {"-alpha", true, -1.0}};
  } else if (lower(algo) == "nonbondedcutoff") {
    if (swfs.size() != names.size()) {
      err = "force '" + fs.id + "': one -switchingFunction per potential expected";
      return false;
    }
    id = "";
    for (const auto& nm : names) id += (id.empty() ? "" : " ") + nm;
    id += " -algorithm NonbondedCutoff";
    for (const auto& s : swfs) id += " -switchingFunction " + s;
    for (size_t i = 0; i < names.size(); i++) {
      if (lower(names[i]) == "coulombewaldreal") params.push_back({"-alpha", false, 0});
      const std::string s = lower(swfs[i]);
      if (s == "c2") params.push_back({"-switchon", false, 0});
      params.push_back({"-cutoff", false, 0});
    }
    params.push_back({"-cutoff", false, 0});
  } else {
    err = "force '" + fs.id + "': algorithm '" + algo + "' is not implemented on the GPU (NonbondedCutoff, PMEwald)";
    return false;
  }
  size_t pos = 0;
  bool found_last = false;
  for (const Param& p : params) {
    bool found = false;
    std::string val;
    if (!p.key.empty()) {
      for (int pass = 0; pass < 2 && !found; pass++) {
        for (size_t j = pass ? 0 : pos; j < pw.size(); j++)
          if (lower(pw[j]) == lower(p.key) && j + 1 < pw.size()) {
            val = pw[j + 1];
            pos = j + 2;
            found = true;
            break;
          }
      }
    } else if (found_last && pos < pw.size()) {
      val = pw[pos++];
      found = true;
    }
    found_last = found;
    double v = 0;
    if (found) {
      if (!to_real(val, v)) {
        err = "force '" + fs.id + "': could not parse '" + val + "' for " + (p.key.empty() ? "a value" : p.key);
        return false;
      }
    } else if (p.has_default) {
      v = p.def;
    } else {
      err = "force '" + fs.id + "': " + (p.key.empty() ? std::string("a value") : p.key) + " is not defined";
      return false;
    }
    values.push_back(v);
  }
  return true;
}

// Molecule::water of buildMoleculeTable (buildTopology.cpp:886-899): three
// atoms whose symbols, in atom order, are H H O, H O H or O H H.
std::vector<char> water_atoms(const System& s) {
  std::vector<std::vector<int>> mol(s.nmolecules);
  for (int i = 0; i < s.n; i++) mol[s.molecule[i]].push_back(i);
  std::vector<char> w(s.n, 0);
  for (const auto& m : mol) {
    if (m.size() != 3) continue;
    std::string p;
    for (int a : m) p += symbol_of(s.type_name[s.type[a]]) == "H" ? "H" : symbol_of(s.type_name[s.type[a]]) == "O" ? "O" : "?";
    if (p == "HHO" || p == "HOH" || p == "OHH")
      for (int a : m) w[a] = 1;
  }
  return w;
}

namespace {

struct V3 { double x, y, z; };
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 at(const std::vector<double>& v, int i) { return {v[3 * i], v[3 * i + 1], v[3 * i + 2]}; }

// Kahan-compensated sum over the atoms as topologyutilities.cpp sums
// momenta: atom 0 starts the sum, atoms 1.. that pass keep() are added
template <class T, class K>
V3 kahan(int n, T term, K keep) {
  V3 sum = term(0), c = {0, 0, 0};
  for (int i = 1; i < n; i++) {
    if (!keep(i)) continue;
    const V3 y = term(i) - c;
    const V3 t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}

}  // namespace

// removeLinearMomentum / removeAngularMomentum (topologyutilities.cpp
// 463-496, 636-675): of the solute (atoms outside water molecules).
void remove_momentum(const System& s, const std::vector<double>& x, std::vector<double>& v, bool linear,
                     bool angular) {
  const int n = s.n;
  if (n == 0) return;
  const std::vector<char> water = water_atoms(s);
  auto solute = [&](int i) { return !water[i]; };
  if (linear) {
    const V3 p = kahan(n, [&](int i) { return at(v, i) * s.mass[i]; }, solute);
    int ns = 0;
    for (int i = 0; i < n; i++) ns += solute(i) ? 1 : 0;
    const V3 avg = p / (double)ns;
    for (int i = 0; i < n; i++)
      if (solute(i)) {
        const V3 d = avg / s.mass[i];
        v[3 * i] -= d.x; v[3 * i + 1] -= d.y; v[3 * i + 2] -= d.z;
      }
  }
  if (angular) {
    // centre of mass of all atoms (centerOfMass, :498-525)
    double sm = s.mass[0];
    V3 com = at(x, 0) * s.mass[0], c = {0, 0, 0};
    for (int i = 1; i < n; i++) {
      sm += s.mass[i];
      const V3 y = at(x, i) * s.mass[i] - c;
      const V3 t = com + y;
      c = (t - com) - y;
      com = t;
    }
    com = com / sm;
    const V3 L = kahan(n, [&](int i) { return cross(at(x, i) - com, at(v, i)) * s.mass[i]; }, solute);
    if (L.x * L.x + L.y * L.y + L.z * L.z == 0) return;
    double I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
      if (!solute(i)) continue;
      const V3 r = at(x, i) - com;
      const double m = s.mass[i];
      const double t[9] = {r.y * r.y + r.z * r.z, -r.x * r.y, -r.x * r.z, -r.x * r.y, r.x * r.x + r.z * r.z,
                           -r.y * r.z, -r.x * r.z, -r.y * r.z, r.x * r.x + r.y * r.y};
      for (int k = 0; k < 9; k++) I[k] += t[k] * m;
    }
    // Matrix3by3::invert (base/Matrix3by3.cpp:490-540): adjugate over the determinant
    double a[9];
    a[0] = I[4] * I[8] - I[5] * I[7];
    a[3] = I[3] * I[8] - I[5] * I[6];
    a[6] = I[3] * I[7] - I[4] * I[6];
    a[1] = I[1] * I[8] - I[2] * I[7];
    a[4] = I[0] * I[8] - I[2] * I[6];
    a[7] = I[0] * I[7] - I[1] * I[6];
    a[2] = I[1] * I[5] - I[2] * I[4];
    a[5] = I[0] * I[5] - I[2] * I[3];
    a[8] = I[0] * I[4] - I[1] * I[3];
    double det = I[0] * a[0] - I[1] * a[3] + I[2] * a[6];
    if (std::fabs(det) < 1e-14 * 1e-14) return;
    det = 1 / det;
    a[0] *= det; a[2] *= det; a[4] *= det; a[6] *= det; a[8] *= det;
    det = -det;
    a[1] *= det; a[3] *= det; a[5] *= det; a[7] *= det;
    const V3 w = {a[0] * L.x + a[1] * L.y + a[2] * L.z, a[3] * L.x + a[4] * L.y + a[5] * L.z,
                  a[6] * L.x + a[7] * L.y + a[8] * L.z};
    for (int i = 0; i < n; i++)
      if (solute(i)) {
        const V3 d = cross(w, at(x, i) - com);
        v[3 * i] -= d.x; v[3 * i + 1] -= d.y; v[3 * i + 2] -= d.z;
      }
  }
}

// randomVelocity (topologyutilities.cpp:24-58): per atom and component
// sqrt(kT / m) times a normal deviate from randomGaussianNumber
// (mathutilities.cpp:186-230, the polar Box-Muller method on drand48 seeded
// by srand48(seed), mathutilities.cpp:147-163, the second deviate of each
// pair kept for the next call).  The generator state is local to this call;
// the reference seeds its process-wide generator on its first call.
std::vector<double> random_velocity(const System& s, double temperature, unsigned int seed) {
  drand48_data st;
  srand48_r((long)seed, &st);
  bool have = false;
  double keep = 0.0;
  auto uniform = [&]() {
    double r;
    drand48_r(&st, &r);
    return r;
  };
  auto gauss = [&]() {
    if (have) {
      have = false;
      return keep;
    }
    double v1 = 0, v2 = 0, r = 2.0;
    while (r >= 1.0 || r < 1.523e-8) {
      v1 = 2.0 * uniform() - 1.0;
      v2 = 2.0 * uniform() - 1.0;
      r = v1 * v1 + v2 * v2;
    }
    const double fac = std::sqrt(-2.0 * std::log(r) / r);
    keep = v1 * fac;
    have = true;
    return v2 * fac;
  };
  const double kbT = temperature * 0.001987191;          // Constant::BOLTZMANN
  std::vector<double> v(3 * (size_t)s.n);
  for (int i = 0; i < s.n; i++) {
    const double f = std::sqrt(kbT / s.mass[i]);
    for (int c = 0; c < 3; c++) v[3 * (size_t)i + c] = f * gauss();
  }
  return v;
}

}  // namespace pmf
