// pmf_build.cpp -- the topology a PSF and a PAR describe.
//
// Restates frontend/buildTopology.cpp:51-727 (atom types, scaled charges,
// bonded parameters with the PAR look-up order and wildcards, the
// Lennard-Jones type table with NBFIX overrides), buildMoleculeTable
// (:729-931, as connected components numbered by their lowest atom) and
// buildExclusionTable (:955-1011, the adds in the same order; the device
// library resolves repeats as ExclusionTable::add does, last add wins).
#include "pmf.h"

#include <cmath>
#include <map>
#include <numeric>

namespace pmf {

namespace {

const double kSqrtCoulomb = 18.2226123264;      // Constant::SQRTCOULOMBCONSTANT
const double kPi = 3.14159265358979323846;      // Constant::M_PI

double deg2rad(double d) { return d * kPi / 180.0; }   // mathutilities.h dtor

// x^6 and x^12 as power<6> / power<12> evaluate them: ((x*x)*x)^2, then squared
double pow6(double x) {
  const double x3 = x * x * x;
  return x3 * x3;
}
double pow12(double x) {
  const double x6 = pow6(x);
  return x6 * x6;
}

// stringutilities.cpp equalWildcard: 2 for an exact match, 1 for a match
// through the wildcards * (any run), % (one character), # (a run of
// digits), + (one digit), 0 otherwise.
int wildcard_match(const std::string& w, const std::string& name) {
  if (w == name) return 2;
  if (w.find_first_of("*%#+") == std::string::npos) return 0;
  size_t pos = 0;
  for (size_t i = 0; i < w.size(); i++) {
    pos = i;
    const char c = w[i];
    if (c == '*' || c == '%' || c == '#' || c == '+') break;
    if (i >= name.size() || c != name[i]) return 0;
  }
  const char c = w[pos];
  const std::string wrest = w.substr(pos + 1);
  auto tail = [&](size_t from) { return from <= name.size() ? name.substr(from) : std::string(); };
  if (pos + 1 == w.size()) {
    if (c == '*') return 1;
    if (c == '#') {
      for (size_t i = pos; i < name.size(); i++)
        if (!std::isdigit((unsigned char)name[i])) return 0;
      return 1;
    }
    if (c == '%' && pos + 1 == name.size()) return 1;
    if (c == '+' && pos + 1 == name.size() && std::isdigit((unsigned char)name[pos])) return 1;
    return 0;
  }
  if (c == '*') {
    int ok = 0;
    for (size_t i = pos; i <= name.size(); i++)
      if (wildcard_match(wrest, tail(i)) > 0) ok = 1;
    return ok;
  }
  if (c == '%') return (pos < name.size() && wildcard_match(wrest, tail(pos + 1)) > 0) ? 1 : 0;
  if (c == '#') {
    int ok = 0;
    for (size_t i = pos; i <= name.size(); i++) {
      if (wildcard_match(wrest, tail(i)) > 0) ok = 1;
      if (i < name.size() && !std::isdigit((unsigned char)name[i])) break;
    }
    return ok;
  }
  if (c == '+')
    return (pos < name.size() && std::isdigit((unsigned char)name[pos]) && wildcard_match(wrest, tail(pos + 1)) > 0)
               ? 1 : 0;
  return 0;
}

// the first best of count candidates graded by wildcard_match (an exact
// match ends the search); -1 when none matches
template <class G>
int best_match(size_t count, G grade_of) {
  int best = -1, grade = 0;
  for (size_t k = 0; k < count; k++) {
    const int g = grade_of(k);
    if (g > grade) {
      best = (int)k;
      grade = g;
    }
    if (grade > 1) break;
  }
  return best;
}

std::string join(std::initializer_list<std::string> parts) {
  std::string s;
  for (const auto& p : parts) s += (s.empty() ? "" : ",") + p;
  return s;
}

}  // namespace

bool parse_exclude(const std::string& s, int& out) {
  const std::string u = upper(trim(s));
  if (u == "NONE") out = EXCL_NONE;
  else if (u == "1-2") out = EXCL_12;
  else if (u == "1-3") out = EXCL_13;
  else if (u == "1-4") out = EXCL_14;
  else if (u == "SCALED1-4") out = EXCL_SCALED14;
  else return false;
  return true;
}

bool build_system(const Psf& psf, const Par& par, int exclude, bool dihedral_mult_psf, System& s, std::string& err) {
  s = System();
  const int n = (int)psf.atoms.size();
  s.n = n;
  s.exclude = exclude;
  std::map<std::string, int> type_of;
  for (const PsfAtom& a : psf.atoms) {
    auto it = type_of.find(a.type);
    int t;
    if (it == type_of.end()) {
      t = (int)s.type_name.size();
      type_of[a.type] = t;
      s.type_name.push_back(a.type);
      s.type_mass.push_back(a.mass);
    } else {
      t = it->second;
    }
    s.type.push_back(t);
    s.charge.push_back(a.charge);
    s.q.push_back(a.charge * kSqrtCoulomb);
    s.mass.push_back(a.mass);
    s.atom_name.push_back(a.name);
    s.resname.push_back(a.resname);
    s.resseq.push_back(a.resseq);
  }
  s.dof = 3 * n - 3;
  auto atom = [&](int one_based, int& out, const char* what) {
    out = one_based - 1;
    if (out < 0 || out >= n) {
      err = std::string("PSF ") + what + " names atom " + std::to_string(one_based) + " of " + std::to_string(n);
      return false;
    }
    return true;
  };
  auto tn = [&](int i) -> const std::string& { return s.type_name[s.type[i]]; };

  // bonds: the PAR entry t1,t2 or t2,t1; the last PAR line of a key wins
  std::map<std::string, const ParBond*> bonds;
  for (const auto& b : par.bonds) bonds[join({b.t[0], b.t[1]})] = &b;
  for (const auto& pb : psf.bonds) {
    int a1, a2;
    if (!atom(pb[0], a1, "bond") || !atom(pb[1], a2, "bond")) return false;
    auto it = bonds.find(join({tn(a1), tn(a2)}));
    if (it == bonds.end()) it = bonds.find(join({tn(a2), tn(a1)}));
    if (it == bonds.end()) {
      err = "could not find bond '" + tn(a1) + "'-'" + tn(a2) + "' (" + std::to_string(pb[0]) + "," +
            std::to_string(pb[1]) + ")";
      return false;
    }
    s.bond_a.push_back(a1); s.bond_b.push_back(a2);
    s.bond_k.push_back(it->second->k); s.bond_r0.push_back(it->second->r0);
  }

  std::map<std::string, const ParAngle*> angles;
  for (const auto& a : par.angles) angles[join({a.t[0], a.t[1], a.t[2]})] = &a;
  for (const auto& pa : psf.angles) {
    int a1, a2, a3;
    if (!atom(pa[0], a1, "angle") || !atom(pa[1], a2, "angle") || !atom(pa[2], a3, "angle")) return false;
    auto it = angles.find(join({tn(a1), tn(a2), tn(a3)}));
    if (it == angles.end()) it = angles.find(join({tn(a3), tn(a2), tn(a1)}));
    if (it == angles.end()) {
      err = "could not find angle '" + tn(a1) + "'-'" + tn(a2) + "'-'" + tn(a3) + "'";
      return false;
    }
    const ParAngle& p = *it->second;
    s.angle_a.push_back(a1); s.angle_b.push_back(a2); s.angle_c.push_back(a3);
    s.angle_k.push_back(p.k);
    s.angle_t0.push_back(deg2rad(p.theta0));
    s.angle_kub.push_back(p.ub ? p.kub : 0.0);
    s.angle_rub.push_back(p.ub ? p.rub : 0.0);
  }

  // dihedrals: exact either way round, then X,b,c,X either way round.  A PSF
  // line repeating the previous dihedral's atoms adds the next PAR term when
  // the multiplicity comes from the PSF.
  std::map<std::string, const ParTorsion*> dih;
  for (const auto& d : par.dihedrals) dih[join({d.t[0], d.t[1], d.t[2], d.t[3]})] = &d;
  for (const auto& pd : psf.dihedrals) {
    int a[4];
    for (int k = 0; k < 4; k++)
      if (!atom(pd[k], a[k], "dihedral")) return false;
    const std::string t1 = tn(a[0]), t2 = tn(a[1]), t3 = tn(a[2]), t4 = tn(a[3]);
    auto it = dih.find(join({t1, t2, t3, t4}));
    if (it == dih.end()) it = dih.find(join({t4, t3, t2, t1}));
    if (it == dih.end()) it = dih.find(join({"X", t2, t3, "X"}));
    if (it == dih.end()) it = dih.find(join({"X", t3, t2, "X"}));
    if (it == dih.end()) {
      err = "could not find dihedral '" + t1 + "'-'" + t2 + "'-'" + t3 + "'-'" + t4 + "'";
      return false;
    }
    const ParTorsion& p = *it->second;
    const bool repeat = !s.dihedrals.empty() && s.dihedrals.back().a[0] == a[0] && s.dihedrals.back().a[1] == a[1] &&
                        s.dihedrals.back().a[2] == a[2] && s.dihedrals.back().a[3] == a[3];
    if (!repeat) {
      Torsion t;
      for (int k = 0; k < 4; k++) t.a[k] = a[k];
      const size_t m = dihedral_mult_psf ? 1 : p.k.size();
      for (size_t u = 0; u < m; u++) {
        t.k.push_back(p.k[u]);
        t.n.push_back(p.n[u]);
        t.phase.push_back(deg2rad(p.phase[u]));
      }
      s.dihedrals.push_back(t);
    } else if (dihedral_mult_psf) {
      Torsion& t = s.dihedrals.back();
      const size_t u = t.k.size();
      if (u >= p.k.size()) {
        err = "PSF multiplicity of dihedral (" + t1 + "," + t2 + "," + t3 + "," + t4 + ") exceeds the PAR definition";
        return false;
      }
      t.k.push_back(p.k[u]);
      t.n.push_back(p.n[u]);
      t.phase.push_back(deg2rad(p.phase[u]));
    } else {
      err = "unexpected PSF multiplicity of dihedral (" + t1 + "," + t2 + "," + t3 + "," + t4 +
            "): use the dihedral multiplicity of the PSF";
      return false;
    }
  }

  // impropers: a,b,c,d; a,X,X,d; X,b,c,d; X,b,c,X; X,X,c,d -- each also reversed
  std::map<std::string, const ParTorsion*> imp;
  for (const auto& d : par.impropers) imp[join({d.t[0], d.t[1], d.t[2], d.t[3]})] = &d;
  for (const auto& pi : psf.impropers) {
    int a[4];
    for (int k = 0; k < 4; k++)
      if (!atom(pi[k], a[k], "improper")) return false;
    const std::string t1 = tn(a[0]), t2 = tn(a[1]), t3 = tn(a[2]), t4 = tn(a[3]);
    const std::string keys[10] = {join({t1, t2, t3, t4}), join({t4, t3, t2, t1}), join({t1, "X", "X", t4}),
                                  join({t4, "X", "X", t1}), join({"X", t2, t3, t4}), join({t4, t3, t2, "X"}),
                                  join({"X", t2, t3, "X"}), join({"X", t3, t2, "X"}), join({"X", "X", t3, t4}),
                                  join({t4, t3, "X", "X"})};
    const ParTorsion* p = nullptr;
    for (const auto& k : keys) {
      auto it = imp.find(k);
      if (it != imp.end()) {
        p = it->second;
        break;
      }
    }
    if (!p) {
      err = "could not find improper '" + t1 + "'-'" + t2 + "'-'" + t3 + "'-'" + t4 + "'";
      return false;
    }
    Torsion t;
    for (int k = 0; k < 4; k++) t.a[k] = a[k];
    t.k.push_back(p->k[0]);
    t.n.push_back(p->n[0]);
    t.phase.push_back(deg2rad(p->phase[0]));
    s.impropers.push_back(t);
  }

  // Lennard-Jones type table: CHARMM combination of the best-matching
  // NONBONDED entries, then the NBFIX pairs over it
  const int nt = (int)s.type_name.size();
  s.ljA.assign((size_t)nt * nt, 0.0);
  s.ljB = s.ljA14 = s.ljB14 = s.ljA;
  std::vector<int> nb_of(nt);
  for (int t = 0; t < nt; t++) {
    nb_of[t] = best_match(par.nonbonded.size(),
                          [&](size_t k) { return wildcard_match(par.nonbonded[k].type, s.type_name[t]); });
    if (nb_of[t] < 0) {
      err = "could not find a nonbonded parameter for atom type '" + s.type_name[t] + "'";
      return false;
    }
  }
  auto set_lj = [&](int i, int j, double A, double B, double A14, double B14) {
    for (int r = 0; r < 2; r++) {
      const size_t k = r ? (size_t)j * nt + i : (size_t)i * nt + j;
      s.ljA[k] = A; s.ljB[k] = B; s.ljA14[k] = A14; s.ljB14[k] = B14;
    }
  };
  for (int i = 0; i < nt; i++)
    for (int j = i; j < nt; j++) {
      const ParNonbonded& p = par.nonbonded[nb_of[i]];
      const ParNonbonded& q = par.nonbonded[nb_of[j]];
      const double r = p.sigma + q.sigma, e = std::sqrt(p.eps * q.eps);
      const double r14 = p.sigma14 + q.sigma14, e14 = std::sqrt(p.eps14 * q.eps14);
      set_lj(i, j, pow12(r) * e, 2 * pow6(r) * e, pow12(r14) * e14, 2 * pow6(r14) * e14);
    }
  for (const ParNbfix& x : par.nbfix) {
    const int bi = best_match(nt, [&](size_t k) { return wildcard_match(x.t[0], s.type_name[k]); });
    if (bi < 0) continue;
    const int bj = best_match(nt, [&](size_t k) { return wildcard_match(x.t[1], s.type_name[k]); });
    if (bj < 0) {
      err = "could not find the types of NBFIX '" + x.t[0] + "' - '" + x.t[1] + "'";
      return false;
    }
    set_lj(bi, bj, x.a, x.b, x.a14, x.b14);
  }

  // molecules: connected components of the bond, angle, dihedral and
  // improper graph, numbered in order of their lowest atom
  std::vector<int> root(n);
  std::iota(root.begin(), root.end(), 0);
  auto find = [&](int x) {
    while (root[x] != x) x = root[x] = root[root[x]];
    return x;
  };
  auto unite = [&](int x, int y) {
    x = find(x); y = find(y);
    if (x != y) root[std::max(x, y)] = std::min(x, y);
  };
  for (size_t k = 0; k < s.bond_a.size(); k++) unite(s.bond_a[k], s.bond_b[k]);
  for (size_t k = 0; k < s.angle_a.size(); k++) { unite(s.angle_a[k], s.angle_b[k]); unite(s.angle_b[k], s.angle_c[k]); }
  for (const auto* list : {&s.dihedrals, &s.impropers})
    for (const Torsion& t : *list) { unite(t.a[0], t.a[1]); unite(t.a[1], t.a[2]); unite(t.a[2], t.a[3]); }
  s.molecule.assign(n, -1);
  std::vector<int> id_of(n, -1);
  for (int i = 0; i < n; i++) {
    const int r = find(i);
    if (id_of[r] < 0) id_of[r] = s.nmolecules++;
    s.molecule[i] = id_of[r];
  }

  // exclusions, in the reference's add order
  if (exclude != EXCL_NONE) {
    auto add = [&](int a, int b, int cls) {
      if (a == b) return;
      s.ex_a.push_back(std::min(a, b));
      s.ex_b.push_back(std::max(a, b));
      s.ex_cls.push_back(cls);
    };
    for (size_t k = 0; k < s.bond_a.size(); k++) add(s.bond_a[k], s.bond_b[k], 2);
    if (exclude != EXCL_12) {
      for (size_t k = 0; k < s.angle_a.size(); k++) add(s.angle_a[k], s.angle_c[k], 2);
      if (exclude != EXCL_13)
        for (const Torsion& t : s.dihedrals) add(t.a[0], t.a[3], exclude == EXCL_14 ? 2 : 1);
    }
  }
  return true;
}

}  // namespace pmf
