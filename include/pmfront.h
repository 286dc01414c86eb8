/* pmfront.h -- C-ABI of the host front end (libpmfront.so): ProtoMol's input
 * files read into the arrays that pmgpu.h takes, and its output formats.
 *
 * Replaces the reference's input and output layer: io/PSFReader.cpp,
 * io/PARReader.cpp, io/PDBReader.cpp, io/XYZReader.cpp,
 * frontend/buildTopology.cpp (topology from PSF + PAR),
 * frontend/OutputEnergies.cpp, io/XYZWriter.cpp, io/PDBWriter.cpp and
 * io/DCDTrajectoryWriter.cpp.  Host code only; nothing here touches a GPU.
 *
 * Every call returns 0 on success or a negative PMF_ERR_* code; the message
 * of the last failure on the calling thread is pmf_last_error().
 */
#ifndef PMFRONT_H
#define PMFRONT_H

#ifdef __cplusplus
extern "C" {
#endif

#define PMF_OK 0
#define PMF_ERR_ARGUMENT (-1)
#define PMF_ERR_IO (-2)
#define PMF_ERR_PARSE (-3)
#define PMF_ERR_TOPOLOGY (-4)

typedef struct pmf_system pmf_system;

const char* pmf_last_error(void);

/* Read a PSF and a PAR file and build the topology the way the reference's
 * protomol application does (applications/protomol-app/protomol.cpp:340-403:
 * PSFReader, PARReader, buildTopology).  exclude is the .conf "exclude" value
 * ("none", "1-2", "1-3", "1-4", "scaled1-4"); dihedral_mult_psf is 0, 1, or -1
 * for the reference's default (PSF multiplicity unless the PAR file is
 * CHARMM22+, protomol.cpp:359-360). */
int pmf_load_system(const char* psf_path, const char* par_path, const char* exclude, int dihedral_mult_psf,
                    pmf_system** out);
void pmf_free_system(pmf_system* s);

/* Sizes: info[PMF_INFO_*] */
enum {
  PMF_INFO_ATOMS = 0,
  PMF_INFO_TYPES,
  PMF_INFO_BONDS,
  PMF_INFO_ANGLES,
  PMF_INFO_DIHEDRALS,
  PMF_INFO_DIHEDRAL_TERMS,
  PMF_INFO_IMPROPERS,
  PMF_INFO_IMPROPER_TERMS,
  PMF_INFO_EXCLUSIONS,
  PMF_INFO_MOLECULES,
  PMF_INFO_DOF,
  PMF_INFO_PAR_DIALECT,        /* 19 or 28: the dialect the PAR file was read in */
  PMF_INFO_DIHEDRAL_MULT_PSF,
  PMF_INFO_COUNT
};
int pmf_system_info(const pmf_system* s, int* info);

/* Per-atom arrays (n entries each; any pointer may be NULL): scaled charge
 * (e * sqrt(332.0636), Atom::scaledCharge), mass, type index, molecule index. */
int pmf_system_atoms(const pmf_system* s, double* q, double* mass, int* type, int* molecule);
/* Lennard-Jones type table, ntypes x ntypes row-major (LennardJonesParameterTable):
 * A, B and the 1-4 values A14, B14. */
int pmf_system_lj(const pmf_system* s, double* A, double* B, double* A14, double* B14);
/* Bond / angle parameters as GenericTopology::bonds / ::angles hold them
 * (rest angle in radians). */
int pmf_system_bonds(const pmf_system* s, int* a, int* b, double* k, double* r0);
int pmf_system_angles(const pmf_system* s, int* a, int* b, int* c, double* k, double* theta0, double* kub,
                      double* rub);
/* Dihedrals (improper = 0) or impropers (improper = 1) as Torsion records:
 * atoms[4 * i + 0..3], terms term_row[i] .. term_row[i + 1] - 1 with force
 * constant, periodicity and phase shift (radians). */
int pmf_system_torsions(const pmf_system* s, int improper, int* atoms, int* term_row, double* k, int* period,
                        double* phase);
/* Exclusion adds in buildExclusionTable order: pair (a < b) and class
 * (1 modified, 2 full); a pair added twice takes the last class. */
int pmf_system_exclusions(const pmf_system* s, int* a, int* b, int* cls);
/* Name of atom type t (PSF type column), NUL-terminated into buf[cap]. */
int pmf_system_type_name(const pmf_system* s, int t, char* buf, int cap);

/* Coordinates of a PDB or XYZ file (InputPosVel: PDB when the file holds an
 * ATOM/HETATM record): *n is the atom count on return; xyz (3 * *n doubles)
 * may be NULL to ask for the count. cap is the capacity of xyz in atoms. */
int pmf_read_coords(const char* path, double* xyz, int cap, int* n);

/* One line of an allEnergies file (frontend/OutputEnergies.cpp:70-109), with
 * its newline: time (fs), E_potential, E_kinetic, E_total, temperature,
 * E_bond, E_angle, E_dihedral, E_improper, E_VdW, E_coulomb, E_other, volume
 * -- values[13] in that order.  Writes at most cap bytes including the NUL;
 * returns the line length (without NUL) or a negative code. */
int pmf_energies_line(const double* values, char* buf, int cap);

/* XYZ coordinate file (io/XYZWriter.cpp) with the system's atom type names,
 * as OutputFinalXYZPos / OutputFinalXYZVel write it. */
int pmf_write_xyz(const char* path, const pmf_system* s, const double* xyz, int n, const char* comment);
/* PDB coordinate file (io/PDBWriter.cpp): the ATOM / HETATM records of the
 * PDB file template_pdb (the position file, as OutputFinalPDBPos uses it)
 * with the coordinates xyz. */
int pmf_write_pdb(const char* path, const char* template_pdb, const double* xyz, int n, const char* comment);

/* DCD trajectory (io/DCDTrajectoryWriter.cpp): nframes frames of n atoms
 * (xyz[3 * (f * n + i) + c]) with the header's first step and time step (fs)
 * as the writer takes them (OutputDCDTrajectory uses 1 and 1.0). */
int pmf_write_dcd(const char* path, int n, int nframes, const double* xyz, int first_step, double timestep_fs,
                  const char* comment);

/* A force definition of a .conf Integrator block (the words after "force",
 * e.g. "LennardJones -algorithm NonbondedCutoff -switchingFunction C2
 * -switchon 7 -cutoff 8") as the force id and positional values that
 * pm_force_create takes: the parameter words are matched to the force's
 * parameter list as ForceFactoryDetails::doMake matches them
 * (frontend/ForceFactoryDetails.cpp:48-190), defaults filled in. */
int pmf_force_values(const char* definition, char* id, int id_cap, double* values, int cap, int* nvalues);

/* Start-up momentum removal of the solute (atoms outside 3-atom waters), as
 * removeLinearMomentum / removeAngularMomentum do it
 * (topology/topologyutilities.cpp:463-496, 636-675); v is updated in place. */
int pmf_remove_momentum(const pmf_system* s, const double* x, double* v, int linear, int angular);

/* Velocities at a temperature as randomVelocity draws them
 * (topology/topologyutilities.cpp:24-58, mathutilities.cpp:147-230: polar
 * Box-Muller on drand48 seeded by srand48(seed)). */
int pmf_random_velocity(const pmf_system* s, double temperature, unsigned int seed, double* v);

#ifdef __cplusplus
}
#endif

#endif /* PMFRONT_H */
