// pmf.h -- host front end: ProtoMol's input files, topology builder, output
// writers and .conf reader, in C++ (no device code).
//
// The front end turns the reference's input files into the plain arrays the
// device library takes (pm_set_topology, pm_set_bonded, pm_set_torsions) and
// writes the reference's output formats.  Every reader follows the file
// grammar of the reader it replaces (cited per function); errors come back as
// a message instead of the reference's report-and-exit.
#pragma once

#include <array>
#include <string>
#include <vector>

namespace pmf {

// ------------------------------------------------------------------ inputs
struct PsfAtom {
  int number = 0;
  std::string seg, resname, name, type;
  int resseq = 0;
  double charge = 0, mass = 0;
  int identity = 0;
};

struct Psf {
  std::vector<PsfAtom> atoms;
  std::vector<std::array<int, 2>> bonds;       // 1-based atom numbers, as in the file
  std::vector<std::array<int, 3>> angles;
  std::vector<std::array<int, 4>> dihedrals, impropers;
};

struct ParBond { std::string t[2]; double k = 0, r0 = 0; };
struct ParAngle { std::string t[3]; double k = 0, theta0 = 0; bool ub = false; double kub = 0, rub = 0; };
struct ParTorsion {
  std::string t[4];
  std::vector<double> k;          // one entry per term (the multiplicity)
  std::vector<int> n;
  std::vector<double> phase;      // degrees, as in the file
};
struct ParNonbonded { std::string type; double eps = 0, sigma = 0, eps14 = 0, sigma14 = 0; };
struct ParNbfix { std::string t[2]; double a = 0, b = 0, a14 = 0, b14 = 0; };

enum ParDialect { PAR_AUTO = 0, PAR_CHARMM19 = 19, PAR_CHARMM28 = 28 };

struct Par {
  int dialect = PAR_CHARMM28;     // the dialect the file was read in
  std::vector<ParBond> bonds;
  std::vector<ParAngle> angles;
  std::vector<ParTorsion> dihedrals, impropers;
  std::vector<ParNonbonded> nonbonded;
  std::vector<ParNbfix> nbfix;
  std::vector<std::string> warnings;
};

// The fields of a PDB ATOM / HETATM record (base/PDB.h PDB::Atom).
struct PdbAtom {
  std::string record, name, altloc, resname, chain, icode, segid, element, charge;
  int serial = 0, resseq = 0;
  double occupancy = 0, tempfactor = 0;
};

// Coordinates of a PDB or XYZ file (positions or velocities).
struct Coords {
  std::vector<double> xyz;                 // 3 per atom
  std::vector<std::string> names;          // XYZ: the name column; PDB: atom names
  std::vector<PdbAtom> pdb;                // PDB only: the records, for the PDB writer
  std::string format;                      // "pdb" or "xyz"
};

bool read_psf(const std::string& path, Psf& psf, std::string& err);
bool read_par(const std::string& path, Par& par, int dialect, std::string& err);
bool read_pdb(const std::string& path, Coords& c, std::string& err);
bool read_xyz(const std::string& path, Coords& c, std::string& err);
// PDB when the file holds an ATOM/HETATM record, else XYZ (InputPosVel order)
bool read_coords(const std::string& path, Coords& c, std::string& err);

// ---------------------------------------------------------------- topology
enum Exclude { EXCL_NONE = 1, EXCL_12 = 2, EXCL_13 = 3, EXCL_14 = 4, EXCL_SCALED14 = 5 };
bool parse_exclude(const std::string& s, int& out);

struct Torsion {
  int a[4];
  std::vector<double> k;
  std::vector<int> n;
  std::vector<double> phase;      // radians
};

struct System {
  int n = 0;
  std::vector<std::string> type_name;      // atom types in order of first appearance
  std::vector<double> type_mass;
  std::vector<int> type;                   // per atom
  std::vector<double> q;                   // charge * sqrt(Coulomb constant)
  std::vector<double> charge;              // charge in e, as read
  std::vector<double> mass;
  std::vector<std::string> atom_name, resname;
  std::vector<int> resseq;
  std::vector<int> bond_a, bond_b;
  std::vector<double> bond_k, bond_r0;
  std::vector<int> angle_a, angle_b, angle_c;
  std::vector<double> angle_k, angle_t0, angle_kub, angle_rub;
  std::vector<Torsion> dihedrals, impropers;
  std::vector<double> ljA, ljB, ljA14, ljB14;   // ntypes x ntypes, row-major, symmetric
  std::vector<int> ex_a, ex_b, ex_cls;          // exclusions in add order (0 none, 1 modified, 2 full)
  std::vector<int> molecule;                    // per atom
  int nmolecules = 0;
  int dof = 0;                                  // 3n - 3
  int exclude = EXCL_SCALED14;
};

bool build_system(const Psf& psf, const Par& par, int exclude, bool dihedral_mult_psf, System& s,
                  std::string& err);

// ------------------------------------------------------------------ output
// One line of the allEnergies file (frontend/OutputEnergies.cpp:70-109):
// time (fs) and the 12 columns, and the optional pressure / molecular
// pressure / molecular temperature / shadow columns when flagged.
struct EnergyRow {
  double time_fs = 0, potential = 0, kinetic = 0, total = 0, temperature = 0;
  double bond = 0, angle = 0, dihedral = 0, improper = 0, lj = 0, coulomb = 0, other = 0, volume = 0;
};
std::string energies_line(const EnergyRow& r);
bool write_xyz(const std::string& path, const std::vector<double>& xyz, const std::vector<std::string>& names,
               const std::string& comment, std::string& err);
// io/PDBWriter.cpp: the input PDB's records with new coordinates
bool write_pdb(const std::string& path, const std::vector<double>& xyz, const std::vector<PdbAtom>& atoms,
               const std::string& comment, std::string& err);

// DCD trajectory (io/DCDTrajectoryWriter.cpp): CHARMM header, then one frame
// of float x, y, z blocks per call.
class DcdWriter {
 public:
  bool open(const std::string& path, int natoms, int first_step, int freq, double timestep_fs,
            const std::string& comment, std::string& err);
  bool frame(const std::vector<double>& xyz, std::string& err);
  void close();
  ~DcdWriter() { close(); }
 private:
  void* f_ = nullptr;
  int natoms_ = 0, frames_ = 0;
  std::string path_;
};

// ------------------------------------------------------------------- .conf
struct ForceSpec {
  std::string id;                    // e.g. "LennardJones Coulomb -algorithm NonbondedCutoff ..."
  std::vector<std::string> names;    // the force names before the first option
  std::vector<std::pair<std::string, std::string>> options;   // "-switchon" -> "7.0", in order
};
struct LevelSpec {
  int level = 0;
  std::string integrator;
  std::vector<std::pair<std::string, std::string>> params;    // lower-cased keyword -> value
  std::vector<ForceSpec> forces;
};
struct Config {
  std::string dir;                                             // directory of the .conf file
  std::vector<std::pair<std::string, std::string>> keys;       // lower-cased keyword -> value, last wins
  std::vector<LevelSpec> levels;
  std::vector<std::string> ignored;                           // words that are no keyword
  bool has(const std::string& k) const;
  std::string get(const std::string& k, const std::string& def = "") const;
  double real(const std::string& k, double def) const;
  long integer(const std::string& k, long def) const;
  bool boolean(const std::string& k, bool def) const;
  std::string path(const std::string& k) const;               // a file name (relative to the working directory)
};
bool read_conf(const std::string& path, Config& cfg, std::string& err);
// "--keyword value" overrides from a command line (parseCommandLine.cpp)
bool apply_overrides(Config& cfg, int argc, const char* const* argv, std::string& err);

// -------------------------------------------------------------- start-up
// A force definition of the .conf (ForceSpec::id) as the device library's
// force id and positional parameter values (pm_force_create).
bool force_values(const ForceSpec& fs, std::string& id, std::vector<double>& values, std::string& err);
std::vector<char> water_atoms(const System& s);
void remove_momentum(const System& s, const std::vector<double>& x, std::vector<double>& v, bool linear,
                     bool angular);
std::vector<double> random_velocity(const System& s, double temperature, unsigned int seed);

// ------------------------------------------------------------------ helpers
bool to_real(const std::string& w, double& v);
bool to_int(const std::string& w, int& v);
std::string trim(const std::string& s);
std::string upper(const std::string& s);
bool starts_nocase(const std::string& prefix, const std::string& word);

}  // namespace pmf
