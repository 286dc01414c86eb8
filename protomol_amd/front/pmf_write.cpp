// pmf_write.cpp -- ProtoMol's output formats: the allEnergies line, XYZ and
// PDB coordinate files and the DCD trajectory.
//
// The layouts are the reference writers' (cited per function) so that files
// from either program read the same; free-text banner lines name this
// program instead of the reference's build.
#include "pmf.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <pwd.h>
#include <unistd.h>

namespace pmf {

namespace {

std::string user_name() {
  const passwd* p = getpwuid(getuid());
  return p ? std::string(p->pw_name) : std::to_string(getuid());
}

// stringutilities.cpp toString(Real): 15 significant digits, shortest form
std::string real15(double x) {
  char b[64];
  std::snprintf(b, sizeof b, "%.15g", x);
  return b;
}

std::string right_fill(const std::string& s, size_t n) { return s.size() < n ? s + std::string(n - s.size(), ' ') : s.substr(0, n); }
std::string left_fill(const std::string& s, size_t n) { return s.size() < n ? std::string(n - s.size(), ' ') + s : s.substr(0, n); }

}  // namespace

// frontend/OutputEnergies.cpp:87-135: the time as %14.2f, then the twelve
// energies, temperature and volume as a 14-wide general float with 8
// significant digits, one blank between columns.
std::string energies_line(const EnergyRow& r) {
  const double v[12] = {r.potential, r.kinetic, r.total, r.temperature, r.bond, r.angle,
                        r.dihedral, r.improper, r.lj, r.coulomb, r.other, r.volume};
  char b[64];
  std::snprintf(b, sizeof b, "%14.2f", r.time_fs);
  std::string s = b;
  for (double x : v) {
    std::snprintf(b, sizeof b, " %14.8g", x);
    s += b;
  }
  s += "\n";
  return s;
}

// io/XYZWriter.cpp:47-70: count, banner + comment, "name\t" and three
// right-aligned 24-wide coordinates with 15 significant digits.
bool write_xyz(const std::string& path, const std::vector<double>& xyz, const std::vector<std::string>& names,
               const std::string& comment, std::string& err) {
  const size_t n = xyz.size() / 3;
  if (names.size() != n) {
    err = "write_xyz: " + std::to_string(names.size()) + " names for " + std::to_string(n) + " atoms";
    return false;
  }
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) {
    err = "cannot open '" + path + "' for writing";
    return false;
  }
  std::fprintf(f, "%zu\n!protomol_amd generated this XYZ file by %s. %s\n", n, user_name().c_str(), comment.c_str());
  for (size_t i = 0; i < n; i++)
    std::fprintf(f, "%s\t%24.15g%24.15g%24.15g\n", names[i].c_str(), xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  const bool ok = std::fclose(f) == 0;
  if (!ok) err = "write error on '" + path + "'";
  return ok;
}

// io/PDBWriter.cpp:47-171: an 80-blank line per atom, then the reference's
// sequence of std::string::replace calls at the PDB.h columns (the 4-wide
// residue name replaces 3 columns, so the rest of the record moves one
// column right, as in the reference), then END.
bool write_pdb(const std::string& path, const std::vector<double>& xyz, const std::vector<PdbAtom>& atoms,
               const std::string& comment, std::string& err) {
  const size_t n = xyz.size() / 3;
  if (atoms.size() != n) {
    err = "write_pdb: the PDB records (" + std::to_string(atoms.size()) + ") do not match the " + std::to_string(n) +
          " atoms (a PDB file is written only for positions read from a PDB file)";
    return false;
  }
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) {
    err = "cannot open '" + path + "' for writing";
    return false;
  }
  std::fprintf(f,
               "REMARKS protomol_amd by %s\nREMARKS This .pdb file was created by PDBWriter\n"
               "REMARKS It was not manually assembled\nREMARKS %s\n",
               user_name().c_str(), comment.c_str());
  for (size_t i = 0; i < n; i++) {
    const PdbAtom& a = atoms[i];
    std::string seq = std::to_string(a.resseq);
    if (a.resseq >= 10000) seq = a.resseq < 36000 ? std::string(1, (char)('A' + (a.resseq - 10000) / 1000)) +
                                                         std::to_string(a.resseq % 1000)
                                                   : "-1";
    std::string line(80, ' ');
    line.replace(0, 6, right_fill(a.record, 6));
    line.replace(6, 5, left_fill(std::to_string(a.serial), 5));
    line.replace(12, 4, right_fill(a.name, 4));
    line.replace(16, 1, right_fill(a.altloc, 1));
    line.replace(17, 3, right_fill(a.resname, 4));
    line.replace(21, 1, right_fill(a.chain, 1));
    line.replace(22, 4, left_fill(seq, 4));
    line.replace(26, 1, right_fill(a.icode, 1));
    for (int k = 0; k < 3; k++) line.replace(31 + 8 * k, 7, left_fill(real15(xyz[3 * i + k]), 7));
    line.replace(54, 6, left_fill(real15(a.occupancy), 6));
    line.replace(60, 6, left_fill(real15(a.tempfactor), 6));
    line.replace(72, 4, left_fill(a.segid, 4));
    line.replace(76, 2, left_fill(a.element, 2));
    line.replace(78, 2, left_fill(a.charge, 2));
    std::fprintf(f, "%s\n", line.c_str());
  }
  std::fprintf(f, "END\n");
  const bool ok = std::fclose(f) == 0;
  if (!ok) err = "write error on '" + path + "'";
  return ok;
}

// io/DCDTrajectoryWriter.cpp:54-226, little endian: the 84-byte CORD block
// (sets, first step, steps between sets, sets again, time step in AKMA
// units at offset 44, CHARMM version 24), the 164-byte title block (two
// 80-character remarks), the atom count block; per frame the x, y and z
// blocks of floats.  The set count at offsets 8 and 20 follows the frames.
bool DcdWriter::open(const std::string& path, int natoms, int first_step, int freq, double timestep_fs,
                     const std::string& comment, std::string& err) {
  close();
  (void)freq;
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    err = "cannot open '" + path + "' for writing";
    return false;
  }
  auto i32 = [&](int32_t v) { std::fwrite(&v, 4, 1, f); };
  auto f32 = [&](float v) { std::fwrite(&v, 4, 1, f); };
  i32(84);
  std::fwrite("CORD", 1, 4, f);
  i32(0);                       // sets, counted as frames are written
  i32(first_step);
  i32(1);
  i32(0);
  for (int k = 0; k < 5; k++) i32(0);
  f32((float)(timestep_fs * 0.02045482828087295));    // Constant::INV_TIMEFACTOR
  for (int k = 0; k < 9; k++) i32(0);
  i32(24);
  i32(84);
  i32(164);
  i32(2);
  const std::string t1 = right_fill("Remarks: File '" + path + "' by " + user_name() + ". protomol_amd", 80);
  const std::string t2 = right_fill("Remarks: " + comment, 80);
  std::fwrite(t1.data(), 1, 80, f);
  std::fwrite(t2.data(), 1, 80, f);
  i32(164);
  i32(4);
  i32(natoms);
  i32(4);
  if (std::ferror(f)) {
    std::fclose(f);
    err = "write error on '" + path + "'";
    return false;
  }
  f_ = f;
  natoms_ = natoms;
  frames_ = 0;
  path_ = path;
  return true;
}

bool DcdWriter::frame(const std::vector<double>& xyz, std::string& err) {
  FILE* f = (FILE*)f_;
  if (!f || xyz.size() != 3 * (size_t)natoms_) {
    err = "DCD frame: writer not open or wrong atom count";
    return false;
  }
  std::vector<float> c(natoms_);
  const int32_t bytes = 4 * natoms_;
  for (int k = 0; k < 3; k++) {
    for (int i = 0; i < natoms_; i++) c[i] = (float)xyz[3 * (size_t)i + k];
    std::fwrite(&bytes, 4, 1, f);
    std::fwrite(c.data(), 4, natoms_, f);
    std::fwrite(&bytes, 4, 1, f);
  }
  frames_++;
  const int32_t sets = frames_;
  std::fseek(f, 8, SEEK_SET);
  std::fwrite(&sets, 4, 1, f);
  std::fseek(f, 20, SEEK_SET);
  std::fwrite(&sets, 4, 1, f);
  std::fseek(f, 0, SEEK_END);
  if (std::ferror(f)) {
    err = "write error on '" + path_ + "'";
    return false;
  }
  std::fflush(f);
  return true;
}

void DcdWriter::close() {
  if (f_) std::fclose((FILE*)f_);
  f_ = nullptr;
}

}  // namespace pmf
