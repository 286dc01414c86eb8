// pmf_api.cpp -- C-ABI of the host front end (include/pmfront.h).
#include "pmfront.h"

#include <cstring>

#include "pmf.h"

struct pmf_system {
  pmf::System s;
  pmf::Par par;
  int dihedral_mult_psf = 0;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <class T>
void put(T* dst, const std::vector<T>& src) {
  if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(T));
}

}  // namespace

extern "C" {

const char* pmf_last_error(void) { return g_err.c_str(); }

int pmf_load_system(const char* psf_path, const char* par_path, const char* exclude, int dihedral_mult_psf,
                    pmf_system** out) {
  if (!psf_path || !par_path || !out) return fail(PMF_ERR_ARGUMENT, "pmf_load_system: null argument");
  *out = nullptr;
  int ex = pmf::EXCL_SCALED14;
  if (exclude && !pmf::parse_exclude(exclude, ex))
    return fail(PMF_ERR_ARGUMENT, std::string("exclude must be none, 1-2, 1-3, 1-4 or scaled1-4 (got '") + exclude + "')");
  std::string err;
  pmf::Psf psf;
  if (!pmf::read_psf(psf_path, psf, err)) return fail(PMF_ERR_PARSE, err);
  auto* h = new pmf_system();
  if (!pmf::read_par(par_path, h->par, pmf::PAR_AUTO, err)) {
    delete h;
    return fail(PMF_ERR_PARSE, err);
  }
  h->dihedral_mult_psf = dihedral_mult_psf < 0 ? (h->par.dialect != pmf::PAR_CHARMM28) : (dihedral_mult_psf != 0);
  if (!pmf::build_system(psf, h->par, ex, h->dihedral_mult_psf != 0, h->s, err)) {
    delete h;
    return fail(PMF_ERR_TOPOLOGY, err);
  }
  *out = h;
  return PMF_OK;
}

void pmf_free_system(pmf_system* s) { delete s; }

int pmf_system_info(const pmf_system* h, int* info) {
  if (!h || !info) return fail(PMF_ERR_ARGUMENT, "pmf_system_info: null argument");
  const pmf::System& s = h->s;
  int dt = 0, it = 0;
  for (const auto& t : s.dihedrals) dt += (int)t.k.size();
  for (const auto& t : s.impropers) it += (int)t.k.size();
  info[PMF_INFO_ATOMS] = s.n;
  info[PMF_INFO_TYPES] = (int)s.type_name.size();
  info[PMF_INFO_BONDS] = (int)s.bond_a.size();
  info[PMF_INFO_ANGLES] = (int)s.angle_a.size();
  info[PMF_INFO_DIHEDRALS] = (int)s.dihedrals.size();
  info[PMF_INFO_DIHEDRAL_TERMS] = dt;
  info[PMF_INFO_IMPROPERS] = (int)s.impropers.size();
  info[PMF_INFO_IMPROPER_TERMS] = it;
  info[PMF_INFO_EXCLUSIONS] = (int)s.ex_a.size();
  info[PMF_INFO_MOLECULES] = s.nmolecules;
  info[PMF_INFO_DOF] = s.dof;
  info[PMF_INFO_PAR_DIALECT] = h->par.dialect;
  info[PMF_INFO_DIHEDRAL_MULT_PSF] = h->dihedral_mult_psf;
  return PMF_OK;
}

int pmf_system_atoms(const pmf_system* h, double* q, double* mass, int* type, int* molecule) {
  if (!h) return fail(PMF_ERR_ARGUMENT, "pmf_system_atoms: null system");
  put(q, h->s.q);
  put(mass, h->s.mass);
  put(type, h->s.type);
  put(molecule, h->s.molecule);
  return PMF_OK;
}

int pmf_system_lj(const pmf_system* h, double* A, double* B, double* A14, double* B14) {
  if (!h) return fail(PMF_ERR_ARGUMENT, "pmf_system_lj: null system");
  put(A, h->s.ljA);
  put(B, h->s.ljB);
  put(A14, h->s.ljA14);
  put(B14, h->s.ljB14);
  return PMF_OK;
}

int pmf_system_bonds(const pmf_system* h, int* a, int* b, double* k, double* r0) {
  if (!h) return fail(PMF_ERR_ARGUMENT, "pmf_system_bonds: null system");
  put(a, h->s.bond_a);
  put(b, h->s.bond_b);
  put(k, h->s.bond_k);
  put(r0, h->s.bond_r0);
  return PMF_OK;
}

int pmf_system_angles(const pmf_system* h, int* a, int* b, int* c, double* k, double* theta0, double* kub,
                      double* rub) {
  if (!h) return fail(PMF_ERR_ARGUMENT, "pmf_system_angles: null system");
  put(a, h->s.angle_a);
  put(b, h->s.angle_b);
  put(c, h->s.angle_c);
  put(k, h->s.angle_k);
  put(theta0, h->s.angle_t0);
  put(kub, h->s.angle_kub);
  put(rub, h->s.angle_rub);
  return PMF_OK;
}

int pmf_system_torsions(const pmf_system* h, int improper, int* atoms, int* term_row, double* k, int* period,
                        double* phase) {
  if (!h) return fail(PMF_ERR_ARGUMENT, "pmf_system_torsions: null system");
  const auto& list = improper ? h->s.impropers : h->s.dihedrals;
  int u = 0;
  for (size_t i = 0; i < list.size(); i++) {
    const pmf::Torsion& t = list[i];
    if (atoms)
      for (int j = 0; j < 4; j++) atoms[4 * i + j] = t.a[j];
    if (term_row) term_row[i] = u;
    for (size_t m = 0; m < t.k.size(); m++, u++) {
      if (k) k[u] = t.k[m];
      if (period) period[u] = t.n[m];
      if (phase) phase[u] = t.phase[m];
    }
  }
  if (term_row) term_row[list.size()] = u;
  return PMF_OK;
}

int pmf_system_exclusions(const pmf_system* h, int* a, int* b, int* cls) {
  if (!h) return fail(PMF_ERR_ARGUMENT, "pmf_system_exclusions: null system");
  put(a, h->s.ex_a);
  put(b, h->s.ex_b);
  put(cls, h->s.ex_cls);
  return PMF_OK;
}

int pmf_system_type_name(const pmf_system* h, int t, char* buf, int cap) {
  if (!h || !buf || cap <= 0) return fail(PMF_ERR_ARGUMENT, "pmf_system_type_name: null argument");
  if (t < 0 || t >= (int)h->s.type_name.size()) return fail(PMF_ERR_ARGUMENT, "pmf_system_type_name: no such type");
  std::strncpy(buf, h->s.type_name[t].c_str(), (size_t)cap - 1);
  buf[cap - 1] = 0;
  return PMF_OK;
}

int pmf_read_coords(const char* path, double* xyz, int cap, int* n) {
  if (!path || !n) return fail(PMF_ERR_ARGUMENT, "pmf_read_coords: null argument");
  pmf::Coords c;
  std::string err;
  if (!pmf::read_coords(path, c, err)) return fail(PMF_ERR_PARSE, err);
  *n = (int)(c.xyz.size() / 3);
  if (xyz) {
    if (cap < *n) return fail(PMF_ERR_ARGUMENT, "pmf_read_coords: buffer holds fewer atoms than the file");
    std::memcpy(xyz, c.xyz.data(), c.xyz.size() * sizeof(double));
  }
  return PMF_OK;
}

int pmf_energies_line(const double* v, char* buf, int cap) {
  if (!v || !buf || cap <= 0) return fail(PMF_ERR_ARGUMENT, "pmf_energies_line: null argument");
  pmf::EnergyRow r;
  r.time_fs = v[0]; r.potential = v[1]; r.kinetic = v[2]; r.total = v[3]; r.temperature = v[4];
  r.bond = v[5]; r.angle = v[6]; r.dihedral = v[7]; r.improper = v[8]; r.lj = v[9]; r.coulomb = v[10];
  r.other = v[11]; r.volume = v[12];
  const std::string s = pmf::energies_line(r);
  if ((int)s.size() >= cap) return fail(PMF_ERR_ARGUMENT, "pmf_energies_line: buffer too small");
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return (int)s.size();
}

int pmf_write_xyz(const char* path, const pmf_system* h, const double* xyz, int n, const char* comment) {
  if (!path || !h || !xyz) return fail(PMF_ERR_ARGUMENT, "pmf_write_xyz: null argument");
  if (n != h->s.n) return fail(PMF_ERR_ARGUMENT, "pmf_write_xyz: atom count differs from the system");
  std::vector<std::string> names(n);
  for (int i = 0; i < n; i++) names[i] = h->s.type_name[h->s.type[i]];
  std::string err;
  if (!pmf::write_xyz(path, std::vector<double>(xyz, xyz + 3 * (size_t)n), names, comment ? comment : "", err))
    return fail(PMF_ERR_IO, err);
  return PMF_OK;
}

int pmf_write_pdb(const char* path, const char* template_pdb, const double* xyz, int n, const char* comment) {
  if (!path || !template_pdb || !xyz) return fail(PMF_ERR_ARGUMENT, "pmf_write_pdb: null argument");
  pmf::Coords c;
  std::string err;
  if (!pmf::read_pdb(template_pdb, c, err)) return fail(PMF_ERR_PARSE, err);
  if (!pmf::write_pdb(path, std::vector<double>(xyz, xyz + 3 * (size_t)n), c.pdb, comment ? comment : "", err))
    return fail(PMF_ERR_IO, err);
  return PMF_OK;
}

int pmf_write_dcd(const char* path, int n, int nframes, const double* xyz, int first_step, double timestep_fs,
                  const char* comment) {
  if (!path || n <= 0 || nframes < 0 || (nframes > 0 && !xyz)) return fail(PMF_ERR_ARGUMENT, "pmf_write_dcd: bad argument");
  pmf::DcdWriter w;
  std::string err;
  if (!w.open(path, n, first_step, 1, timestep_fs, comment ? comment : "", err)) return fail(PMF_ERR_IO, err);
  for (int f = 0; f < nframes; f++)
    if (!w.frame(std::vector<double>(xyz + 3 * (size_t)f * n, xyz + 3 * (size_t)(f + 1) * n), err))
      return fail(PMF_ERR_IO, err);
  w.close();
  return PMF_OK;
}

int pmf_force_values(const char* definition, char* id, int id_cap, double* values, int cap, int* nvalues) {
  if (!definition || !id || id_cap <= 0 || !nvalues) return fail(PMF_ERR_ARGUMENT, "pmf_force_values: null argument");
  pmf::ForceSpec fs;
  fs.id = definition;
  std::string out, err;
  std::vector<double> v;
  if (!pmf::force_values(fs, out, v, err)) return fail(PMF_ERR_PARSE, err);
  if ((int)out.size() >= id_cap || (int)v.size() > cap) return fail(PMF_ERR_ARGUMENT, "pmf_force_values: buffer too small");
  std::memcpy(id, out.c_str(), out.size() + 1);
  if (values && !v.empty()) std::memcpy(values, v.data(), v.size() * sizeof(double));
  *nvalues = (int)v.size();
  return PMF_OK;
}

int pmf_remove_momentum(const pmf_system* h, const double* x, double* v, int linear, int angular) {
  if (!h || !x || !v) return fail(PMF_ERR_ARGUMENT, "pmf_remove_momentum: null argument");
  const size_t m = 3 * (size_t)h->s.n;
  std::vector<double> xv(x, x + m), vv(v, v + m);
  pmf::remove_momentum(h->s, xv, vv, linear != 0, angular != 0);
  std::memcpy(v, vv.data(), m * sizeof(double));
  return PMF_OK;
}

int pmf_random_velocity(const pmf_system* h, double temperature, unsigned int seed, double* v) {
  if (!h || !v) return fail(PMF_ERR_ARGUMENT, "pmf_random_velocity: null argument");
  if (temperature < 0) return fail(PMF_ERR_ARGUMENT, "pmf_random_velocity: temperature < 0");
  const std::vector<double> r = pmf::random_velocity(h->s, temperature, seed);
  std::memcpy(v, r.data(), r.size() * sizeof(double));
  return PMF_OK;
}

}  // extern "C"
