// pmf_read.cpp -- readers of ProtoMol's input files: PSF, PAR (CHARMM19 /
// X-PLOR and CHARMM22/27 dialects), PDB and XYZ.
//
// Each reader keeps the grammar of the reader it replaces (cited per
// function) -- which lines are headers, how records are counted, which
// record shapes are accepted -- because the topology built from the files
// must be the reference's.  The code is a restatement over a whole-file
// line buffer with a token cursor, not the reference's stream loop.
#include "pmf.h"

#include <cctype>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace pmf {

// ----------------------------------------------------------------- helpers
std::string trim(const std::string& s) {
  const char* ws = " \t\n\r";
  size_t a = s.find_first_not_of(ws);
  if (a == std::string::npos) return "";
  return s.substr(a, s.find_last_not_of(ws) + 1 - a);
}

std::string upper(const std::string& s) {
  std::string u(s);
  for (auto& ch : u) ch = (char)std::toupper((unsigned char)ch);
  return u;
}

// word begins with prefix, ignoring case (stringutilities.cpp equalStartNocase)
bool starts_nocase(const std::string& prefix, const std::string& word) {
  if (prefix.size() > word.size()) return false;
  for (size_t i = 0; i < prefix.size(); i++)
    if (std::toupper((unsigned char)prefix[i]) != std::toupper((unsigned char)word[i])) return false;
  return true;
}

static bool blank_tail(const char* p) {
  for (; *p; p++)
    if (!std::isspace((unsigned char)*p)) return false;
  return true;
}

static bool printable(const std::string& w) {
  for (char ch : w)
    if (!std::isprint((unsigned char)ch) && !std::isspace((unsigned char)ch)) return false;
  return true;
}

// a whole token that strtod reads, normal or zero (stringutilities.cpp toReal)
bool to_real(const std::string& w, double& v) {
  if (w.empty()) return false;
  char* end = nullptr;
  errno = 0;
  const double d = std::strtod(w.c_str(), &end);
  v = d;
  const double a = std::fabs(d);
  return ((a >= DBL_MIN && a <= DBL_MAX) || a == 0.0) && errno != ERANGE && end && blank_tail(end) && printable(w);
}

// a whole integer token, or a real with an integral value (stringutilities.cpp toInt)
bool to_int(const std::string& w, int& v) {
  if (w.empty()) return false;
  char* end = nullptr;
  errno = 0;
  const long l = std::strtol(w.c_str(), &end, 10);
  v = (int)l;
  if ((long)v == l && errno != ERANGE && end && blank_tail(end) && printable(w)) return true;
  double r;
  if (to_real(w, r)) {
    v = (int)(unsigned int)r;
    return (double)v == r;
  }
  return false;
}

static double real_or_zero(const std::string& w) {
  double v = 0;
  to_real(w, v);
  return v;
}

static int int_or_zero(const std::string& w) {
  int v = 0;
  to_int(w, v);
  return v;
}

static bool read_lines(const std::string& path, std::vector<std::string>& lines, std::string& err,
                       const char* what) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = std::string("cannot open ") + what + " file '" + path + "'";
    return false;
  }
  std::string l;
  while (std::getline(f, l)) {
    if (!l.empty() && l.back() == '\r') l.pop_back();
    lines.push_back(l);
  }
  return true;
}

// Token cursor over a line buffer: tokens may run across lines, and the rest
// of the line after the last token read is what the next line read returns
// (the behaviour of operator>> followed by getline).
struct Cursor {
  const std::vector<std::string>& L;
  size_t li = 0, ci = 0;
  explicit Cursor(const std::vector<std::string>& l) : L(l) {}
  bool eof() const { return li >= L.size(); }
  bool token(std::string& t) {
    while (li < L.size()) {
      const std::string& s = L[li];
      while (ci < s.size() && std::isspace((unsigned char)s[ci])) ci++;
      if (ci < s.size()) {
        size_t b = ci;
        while (ci < s.size() && !std::isspace((unsigned char)s[ci])) ci++;
        t = s.substr(b, ci - b);
        return true;
      }
      li++;
      ci = 0;
    }
    return false;
  }
  bool line(std::string& out) {
    if (li >= L.size()) return false;
    out = L[li].substr(std::min(ci, L[li].size()));
    li++;
    ci = 0;
    return true;
  }
};

// ---------------------------------------------------------------------- PSF
// io/PSFReader.cpp:53-330: the first word is PSF; a line holding '!' is a
// record header "<count> !<KEYWORD>" (or "<count> <count2> !NGRP"); the
// records that follow are read as whitespace-separated fields.
bool read_psf(const std::string& path, Psf& psf, std::string& err) {
  std::vector<std::string> L;
  if (!read_lines(path, L, err, "PSF")) return false;
  psf = Psf();
  Cursor cur(L);
  std::string head;
  if (!cur.token(head) || upper(head) != "PSF") {
    err = "'" + path + "' is not a PSF file (no PSF header)";
    return false;
  }
  std::string line;
  auto need = [&](std::string& t, const char* rec) {
    if (!cur.token(t)) {
      err = std::string("PSF file '") + path + "' ends inside the " + rec + " records";
      return false;
    }
    return true;
  };
  while (cur.line(line)) {
    line = trim(line);
    if (line.find('!') == std::string::npos) continue;
    std::istringstream ss(line);
    std::vector<std::string> h;
    int index = -1;
    std::string w;
    while (ss >> w) {
      if (w[0] == '!' && index < 0) index = (int)h.size();
      h.push_back(w);
    }
    if (index < 1 || index > 2 || h.size() < 2) {
      err = "PSF file '" + path + "' has a corrupt record header: '" + line + "'";
      return false;
    }
    const std::string& key = h[index];
    int count = int_or_zero(h[0]);
    if (index == 1) {
      int tmp;
      if (!to_int(h[0], tmp)) {
        err = "PSF file '" + path + "': record count is not an integer: '" + line + "'";
        return false;
      }
      std::string t;
      if (starts_nocase("!NTITLE", key)) {
        for (int k = 0; k < count; k++) cur.line(t);
        continue;
      }
      if (starts_nocase("!NATOM", key)) {
        for (int k = 0; k < count; k++) {
          PsfAtom a;
          std::string f[9];
          for (auto& x : f)
            if (!need(x, "atom")) return false;
          a.number = int_or_zero(f[0]);
          a.seg = f[1];
          a.resseq = int_or_zero(f[2]);
          a.resname = f[3];
          a.name = f[4];
          a.type = f[5];
          if (!to_real(f[6], a.charge) || !to_real(f[7], a.mass)) {
            err = "PSF file '" + path + "': atom " + f[0] + " has a bad charge or mass";
            return false;
          }
          a.identity = int_or_zero(f[8]);
          psf.atoms.push_back(a);
        }
        continue;
      }
      auto ints = [&](int per, const char* rec, std::vector<int>& out) {
        out.resize((size_t)count * per);
        for (size_t k = 0; k < out.size(); k++) {
          if (!need(t, rec)) return false;
          if (!to_int(t, out[k])) {
            err = std::string("PSF file '") + path + "': bad atom number '" + t + "' in the " + rec + " records";
            return false;
          }
        }
        return true;
      };
      std::vector<int> v;
      if (starts_nocase("!NBOND", key)) {
        if (!ints(2, "bond", v)) return false;
        for (int k = 0; k < count; k++) psf.bonds.push_back({v[2 * k], v[2 * k + 1]});
        continue;
      }
      if (starts_nocase("!NTHETA", key)) {
        if (!ints(3, "angle", v)) return false;
        for (int k = 0; k < count; k++) psf.angles.push_back({v[3 * k], v[3 * k + 1], v[3 * k + 2]});
        continue;
      }
      if (starts_nocase("!NPHI", key)) {
        if (!ints(4, "dihedral", v)) return false;
        for (int k = 0; k < count; k++) psf.dihedrals.push_back({v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]});
        continue;
      }
      if (starts_nocase("!NIMPHI", key)) {
        if (!ints(4, "improper", v)) return false;
        for (int k = 0; k < count; k++) psf.impropers.push_back({v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]});
        continue;
      }
      if (starts_nocase("!NDON", key) || starts_nocase("!NACC", key)) {
        if (!ints(2, "donor/acceptor", v)) return false;
        continue;
      }
      if (starts_nocase("!NNB", key)) {
        if (!ints(1, "NNB", v)) return false;
        continue;
      }
    } else if (starts_nocase("!NGRP", key)) {
      std::string t;
      for (long k = 0; k < 3L * count; k++)
        if (!need(t, "NGRP")) return false;
      continue;
    }
    // any other record is skipped unread, as the reference skips it
  }
  return true;
}

// ---------------------------------------------------------------------- PAR
namespace {

bool kw19(const std::string& w) {
  for (const char* k : {"bond", "angl", "dihe", "impr", "nonb", "nbfi", "hbon"})
    if (starts_nocase(k, w)) return true;
  return false;
}
bool kw28(const std::string& w) {
  for (const char* k : {"BOND", "ANGL", "THET", "DIHE", "PHI", "IMPH", "IMPROPER", "NBOND", "NONBONDED", "NBFIX",
                        "HBOND", "END"})
    if (starts_nocase(k, w)) return true;
  return false;
}

struct ParRecord {
  std::vector<std::string> tok;
  std::string shape;     // 'd' for a number, 'w' for a word, per token
  int line;
};

enum Section { S_NONE, S_BOND, S_ANGLE, S_DIHEDRAL, S_IMPROPER, S_NONBONDED, S_NBFIX, S_HBOND };

}  // namespace

// io/PARReader.cpp:96-503.  Pass 1 strips titles ('*'), {...} comments and
// '!' comments, splits each line into tokens with a d/w shape, drops
// REMARK/SET lines and votes on the dialect; pass 2 walks the records by
// section and accepts the record shapes the reference accepts.
bool read_par(const std::string& path, Par& par, int dialect, std::string& err) {
  std::vector<std::string> L;
  if (!read_lines(path, L, err, "PAR")) return false;
  par = Par();
  std::vector<ParRecord> recs;
  int v19 = 0, v28 = 0, depth = 0, lineno = 0;
  for (const std::string& raw : L) {
    ++lineno;
    std::string s = trim(raw);
    if (!s.empty() && s[0] == '*') continue;
    const bool brace = s.find('{') != std::string::npos || s.find('}') != std::string::npos;
    if (!brace) {
      if (depth > 0) continue;
    } else {
      std::string kept;
      for (size_t i = 0; i < s.size(); i++) {
        if (s[i] == '{') {
          ++depth;
        } else if (s[i] == '}') {
          --depth;
          kept += ' ';
          if (depth > 0) {
            std::istringstream rest(s.substr(i));
            std::string r;
            rest >> r;
            std::string more;
            if (r == "}" && !(rest >> more)) {
              depth = 0;
              par.warnings.push_back("comment not closed properly at line " + std::to_string(lineno));
            }
          }
          if (depth < 0) {
            par.warnings.push_back("extra '}' at line " + std::to_string(lineno));
            depth = 0;
          }
        } else if (depth == 0) {
          kept += s[i];
        }
      }
      s = kept;
    }
    s = s.substr(0, s.find('!'));
    ParRecord r;
    r.line = lineno;
    std::istringstream ss(s);
    std::string t;
    while (ss >> t) {
      if (t[0] == '*') break;
      double d;
      r.tok.push_back(t);
      r.shape += to_real(t, d) ? 'd' : 'w';
    }
    if (!r.tok.empty() && (starts_nocase("REMARK", r.tok[0]) || starts_nocase("SET", r.tok[0]))) continue;
    if (r.tok.empty()) continue;
    const bool has_word = r.shape.find('w') != std::string::npos;
    if (r.shape.size() == 1 && kw28(r.tok[0]))
      ++v28;
    else if ((r.shape.size() > 1 && kw19(r.tok[0])) || !has_word)
      ++v19;
    else if (r.shape.size() > 1 && !kw19(r.tok[0]) && !kw28(r.tok[0]))
      ++v28;
    recs.push_back(std::move(r));
  }
  const bool c19 = dialect == PAR_AUTO ? v19 > v28 : dialect == PAR_CHARMM19;
  par.dialect = (v19 > v28) ? PAR_CHARMM19 : PAR_CHARMM28;   // the detected dialect
  const double sig19 = std::pow(2.0, -5.0 / 6.0);
  Section sec = S_NONE;
  for (size_t ri = 0; ri < recs.size(); ri++) {
    const ParRecord& r = recs[ri];
    size_t j = 0;
    std::string shape = r.shape;
    const std::string& w0 = r.tok[0];
    const std::string at = " at line " + std::to_string(r.line);
    if (c19) {
      bool skip = false;
      for (const char* k : {"AEXP", "REXP", "HAEX", "AAEX", "NBOND", "CUTNB", "END", "CTONN", "EPS", "VSWI", "NBXM",
                            "INHI"})
        if (starts_nocase(k, w0)) skip = true;
      if (skip) continue;
      if (starts_nocase("bond", w0)) sec = S_BOND;
      else if (starts_nocase("angl", w0)) sec = S_ANGLE;
      else if (starts_nocase("dihe", w0)) sec = S_DIHEDRAL;
      else if (starts_nocase("impr", w0)) sec = S_IMPROPER;
      else if (starts_nocase("nonb", w0)) sec = S_NONBONDED;
      else if (starts_nocase("nbfi", w0)) sec = S_NBFIX;
      else if (starts_nocase("hbon", w0)) sec = S_HBOND;
      else {
        par.warnings.push_back("unknown parameter '" + w0 + "'" + at);
        continue;
      }
      j = 1;
      shape = shape.substr(1);
    } else {
      if (starts_nocase("bond", w0)) { sec = S_BOND; continue; }
      if (starts_nocase("angl", w0) || starts_nocase("thet", w0)) { sec = S_ANGLE; continue; }
      if (starts_nocase("dihe", w0) || starts_nocase("phi", w0)) { sec = S_DIHEDRAL; continue; }
      if (starts_nocase("impr", w0) || starts_nocase("imph", w0)) { sec = S_IMPROPER; continue; }
      if (starts_nocase("nonb", w0) || starts_nocase("nbon", w0)) { sec = S_NONBONDED; continue; }
      if (starts_nocase("nbfi", w0)) { sec = S_NBFIX; continue; }
      if (starts_nocase("hbon", w0)) { sec = S_HBOND; continue; }
      bool opt = false;
      for (const char* k : {"end", "nbxm", "grou", "cdie", "shif", "vgro", "vdis", "vswi", "cutn", "ctof", "cton", "eps",
                            "e14f", "wmin", "aexp", "rexp", "haex", "aaex", "noac", "hbno", "cuth"})
        if (starts_nocase(k, w0)) opt = true;
      if (opt) {
        if (!(sec == S_NONBONDED || sec == S_NBFIX || sec == S_HBOND))
          par.warnings.push_back("unknown parameter '" + w0 + "'" + at);
        continue;
      }
    }
    auto T = [&](size_t k) -> const std::string& { return r.tok[j + k]; };
    auto D = [&](size_t k) { return real_or_zero(T(k)); };
    switch (sec) {
      case S_BOND:
        if (shape == "wwdd") {
          ParBond b;
          b.t[0] = T(0); b.t[1] = T(1); b.k = D(2); b.r0 = D(3);
          par.bonds.push_back(b);
        } else {
          par.warnings.push_back("unknown bond definition (" + shape + ")" + at);
        }
        break;
      case S_ANGLE: {
        ParAngle a;
        a.t[0] = T(0); a.t[1] = T(1); a.t[2] = T(2);
        if (shape == "wwwdd") {
          a.k = D(3); a.theta0 = D(4);
        } else if (shape == "wwwdddd") {
          a.k = D(3); a.theta0 = D(4); a.ub = true; a.kub = D(5); a.rub = D(6);
        } else if (shape == "wwwddwdd") {
          a.k = D(3); a.theta0 = D(4); a.ub = true; a.kub = D(6); a.rub = D(7);
        } else {
          par.warnings.push_back("unknown angle definition (" + shape + ")" + at);
          break;
        }
        par.angles.push_back(a);
        break;
      }
      case S_DIHEDRAL:
        if (shape == "wwwwddd") {
          const std::string t0 = T(0), t1 = T(1), t2 = T(2), t3 = T(3);
          bool extend = false;
          if (!par.dihedrals.empty() && !c19) {
            const ParTorsion& p = par.dihedrals.back();
            extend = (p.t[0] == t0 && p.t[1] == t1 && p.t[2] == t2 && p.t[3] == t3) ||
                     (p.t[0] == t3 && p.t[1] == t2 && p.t[2] == t1 && p.t[3] == t0);
          }
          if (!extend) {
            ParTorsion p;
            p.t[0] = t0; p.t[1] = t1; p.t[2] = t2; p.t[3] = t3;
            par.dihedrals.push_back(p);
          }
          ParTorsion& p = par.dihedrals.back();
          p.k.push_back(D(4));
          p.n.push_back(int_or_zero(T(5)));
          p.phase.push_back(D(6));
        } else if (shape == "wwwwwdddd") {
          // X-PLOR "MULTiple m": the first term here, m-1 more on the lines after
          ParTorsion p;
          p.t[0] = T(0); p.t[1] = T(1); p.t[2] = T(2); p.t[3] = T(3);
          p.k.push_back(D(6)); p.n.push_back(int_or_zero(T(7))); p.phase.push_back(D(8));
          const int mult = int_or_zero(T(5));
          for (int m = 1; m < mult; m++) {
            if (ri + 1 >= recs.size()) {
              par.warnings.push_back("incomplete dihedral definition" + at);
              break;
            }
            const ParRecord& q = recs[ri + 1];
            if (q.shape.compare(0, 3, "ddd") != 0) {
              par.warnings.push_back("wrong multiple dihedral definition at line " + std::to_string(q.line));
              break;
            }
            ri++;
            p.k.push_back(real_or_zero(q.tok[0]));
            p.n.push_back(int_or_zero(q.tok[1]));
            p.phase.push_back(real_or_zero(q.tok[2]));
          }
          par.dihedrals.push_back(p);
        } else {
          par.warnings.push_back("unknown dihedral definition (" + shape + ")" + at);
        }
        break;
      case S_IMPROPER:
        if (shape == "wwwwddd") {
          ParTorsion p;
          p.t[0] = T(0); p.t[1] = T(1); p.t[2] = T(2); p.t[3] = T(3);
          p.k.push_back(D(4)); p.n.push_back(int_or_zero(T(5))); p.phase.push_back(D(6));
          par.impropers.push_back(p);
        } else {
          par.warnings.push_back("unknown improper definition (" + shape + ")" + at);
        }
        break;
      case S_NONBONDED: {
        if ((!c19 && shape != "wdddddd" && shape != "wddd") || (c19 && shape != "wdddd")) {
          par.warnings.push_back("unknown nonbonded definition (" + shape + ")" + at);
          break;
        }
        // CHARMM22+: type, ignored, epsilon, Rmin/2 [, ignored, eps14, Rmin14/2];
        // CHARMM19: type, epsilon, sigma, eps14, sigma14 (sigma -> Rmin/2)
        ParNonbonded nb;
        nb.type = T(0);
        const int o = c19 ? 0 : 1;
        const double f = c19 ? sig19 : 1.0;
        nb.eps = D(1 + o);
        nb.sigma = D(2 + o) * f;
        nb.eps14 = shape.size() >= 5 ? D(3 + 2 * o) : nb.eps;
        nb.sigma14 = (shape.size() >= 5 ? D(4 + 2 * o) : nb.sigma) * f;
        par.nonbonded.push_back(nb);
        break;
      }
      case S_NBFIX:
        if (shape == "wwdddd" || shape == "wwdd") {
          ParNbfix x;
          x.t[0] = T(0); x.t[1] = T(1);
          x.a = D(2); x.b = D(3);
          x.a14 = shape == "wwdddd" ? D(4) : x.a;
          x.b14 = shape == "wwdddd" ? D(5) : x.b;
          par.nbfix.push_back(x);
        } else {
          par.warnings.push_back("unknown nbfix definition (" + shape + ")" + at);
        }
        break;
      case S_HBOND:
        break;
      default:
        par.warnings.push_back("definition outside any section (" + shape + ")" + at);
        break;
    }
  }
  return true;
}

// ---------------------------------------------------------------------- PDB
static bool pdb_has_atoms(const std::vector<std::string>& L) {
  for (const auto& l : L) {
    std::istringstream ss(l);
    std::string w;
    ss >> w;
    if (w == "ATOM" || w == "HETATM") return true;
  }
  return false;
}

static std::string col(const std::string& rec, size_t at, size_t len) { return trim(rec.substr(at, len)); }

// io/PDBReader.cpp:91-189: fixed columns (PDB.h:66-102) of ATOM / HETATM
// records up to the first END; the residue name takes 4 columns (TIP3).
bool read_pdb(const std::string& path, Coords& c, std::string& err) {
  std::vector<std::string> L;
  if (!read_lines(path, L, err, "PDB")) return false;
  if (!pdb_has_atoms(L)) {
    err = "'" + path + "' holds no ATOM or HETATM record";
    return false;
  }
  c = Coords();
  c.format = "pdb";
  for (const auto& l : L) {
    if (l.compare(0, 3, "END") == 0) break;
    if (l.compare(0, 4, "ATOM") != 0 && l.compare(0, 6, "HETATM") != 0) continue;
    std::string rec = l;
    if (rec.size() < 80) rec.resize(80, ' ');
    PdbAtom a;
    a.record = col(rec, 0, 6);
    a.serial = int_or_zero(col(rec, 6, 5));
    a.name = col(rec, 12, 4);
    a.altloc = col(rec, 16, 1);
    a.resname = col(rec, 17, 4);
    a.chain = col(rec, 21, 1);
    const std::string rs = col(rec, 22, 4);
    if (!to_int(rs, a.resseq)) {
      int tail;
      if (rs.size() == 4 && to_int(rs.substr(1), tail) && rs[0] >= 'A' && rs[0] <= 'Z')
        a.resseq = (rs[0] - 'A') * 1000 + 10000 + tail;
      else
        a.resseq = -1;
    }
    a.icode = col(rec, 26, 1);
    a.occupancy = std::strtod(rec.substr(54, 6).c_str(), nullptr);
    a.tempfactor = std::strtod(rec.substr(60, 6).c_str(), nullptr);
    a.segid = col(rec, 72, 4);
    a.element = col(rec, 76, 2);
    a.charge = col(rec, 78, 2);
    c.names.push_back(a.name);
    c.pdb.push_back(a);
    for (size_t k = 0; k < 3; k++) c.xyz.push_back(std::strtod(rec.substr(30 + 8 * k, 8).c_str(), nullptr));
  }
  return true;
}

// ---------------------------------------------------------------------- XYZ
// io/XYZReader.cpp:79-103: the atom count, a comment line, then
// "name x y z" per atom.
bool read_xyz(const std::string& path, Coords& c, std::string& err) {
  std::vector<std::string> L;
  if (!read_lines(path, L, err, "XYZ")) return false;
  c = Coords();
  c.format = "xyz";
  Cursor cur(L);
  std::string t, rest;
  int n = 0;
  if (!cur.token(t) || !to_int(t, n) || n < 0) {
    err = "'" + path + "' is not an XYZ file (no atom count)";
    return false;
  }
  cur.line(rest);          // rest of the count line
  cur.line(rest);          // comment
  c.xyz.resize(3 * (size_t)n);
  c.names.resize(n);
  for (int i = 0; i < n; i++) {
    std::string x[3];
    if (!cur.token(c.names[i]) || !cur.token(x[0]) || !cur.token(x[1]) || !cur.token(x[2])) {
      err = "XYZ file '" + path + "' ends after " + std::to_string(i) + " of " + std::to_string(n) + " atoms";
      return false;
    }
    for (int k = 0; k < 3; k++)
      if (!to_real(x[k], c.xyz[3 * (size_t)i + k])) {
        err = "XYZ file '" + path + "': bad coordinate '" + x[k] + "' of atom " + std::to_string(i + 1);
        return false;
      }
  }
  return true;
}

bool read_coords(const std::string& path, Coords& c, std::string& err) {
  std::vector<std::string> L;
  if (!read_lines(path, L, err, "coordinate")) return false;
  if (pdb_has_atoms(L)) return read_pdb(path, c, err);
  return read_xyz(path, c, err);
}

}  // namespace pmf
