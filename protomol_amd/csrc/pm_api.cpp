// pm_api.cpp -- host side of the C ABI declared in include/pmgpu.h.
//
// Owns the device buffers, restates the reference's host-side set-up
// arithmetic (periodic box, cell grid, enumerator delta list, Ewald alpha,
// PME tables) and queues the gfx950 kernels on one HIP stream per context.
#include "pmgpu.h"
#include "pm_internal.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

// automatic pair-list skin (A) when PM_OPT_LIST_SKIN is -1 (98k water box,
// 1 fs, exact list test, pair kernel with a lane per list entry).  Mixed
// precision: 0.114 builds per step and 386.9 ns/day at 2.2 A, against 384.1
// at 1.8 A, 375.6 at 1.4 A (0.18 builds per step) and 384.5 at 3.0 A.  FP64,
// where an entry costs more (cell 4.0 A, 200 steps): 218.6 / 220.5 / 221.5 /
// 213.6 / 207.9 / 200.5 ns/day at 0.6 / 0.8 / 1.0 / 1.4 / 1.8 / 2.2 A.
#define PM_LIST_SKIN_AUTO_MIXED 2.2
#define PM_LIST_SKIN_AUTO_DOUBLE 1.0
// automatic inner-list skin (A) when PM_OPT_PRUNE_SKIN is -1: none (on the
// 98k water box at 1 fs the fastest hydrogens cross half of a 0.2-0.4 A
// inner skin almost every step, and the prune pass costs more than the
// 20-25 % of pair lanes it saves)
#define PM_PRUNE_SKIN_AUTO 0.0

// launch recorder of the calling thread (replica batches)
static thread_local std::vector<PmRecord>* g_recorder = nullptr;
void pm_set_recorder(void* rec) { g_recorder = static_cast<std::vector<PmRecord>*>(rec); }
bool pm_recording() { return g_recorder != nullptr; }
void pm_record(int kind, int variant, unsigned blocks, unsigned threads, unsigned lds, const void* args,
               unsigned argbytes) {
  if (!g_recorder) return;
  g_recorder->emplace_back();
  PmRecord& r = g_recorder->back();
  r.kind = argbytes <= PM_REC_ARGMAX ? kind : -1;
  r.variant = variant;
  r.blocks = blocks;
  r.threads = threads;
  r.lds = lds;
  r.argbytes = argbytes <= PM_REC_ARGMAX ? argbytes : 0;
  if (args && r.argbytes) std::memcpy(r.args, args, r.argbytes);
}

namespace {

const double kEpsilon = 1.0e-14;                 // base/pmconstants.cpp:45
const int kMaxInt2 = 2147483647 / 2;             // base/pmconstants.cpp:41
const double kPi = 3.14159265358979323846;
const double kInvTimeFactor = 0.02045482828087295;
const double kTimeFactor = 48.88821290839616;   // Constant::TIMEFACTOR (base/pmconstants.cpp:48)

struct Term {
  int pot;      // 0 LJ, 1 Coulomb, 2 CoulombEwaldReal
  int swf;      // PM_SWF_C1, PM_SWF_C2, PM_SWF_CUTOFF
  double alpha = 0, switchon = 0, cutoff = 0;
};

struct ForceDesc {
  bool pme = false;
  int bonded = 0;   // 1 "Bond", 2 "Angle" (forces/BondSystemForce.h, AngleSystemForce.h)
  // nonbonded cutoff
  std::vector<Term> terms;
  double algo_cutoff = 0;
  // PME
  int nx = 0, ny = 0, nz = 0, order = 4;
  double rc = 0, accuracy = 1e-6, alpha_in = -1;
  bool real = false, recip = false, corr = false;
  // cached PME state (NonbondedPMEwaldSystemForce::initialize, :499-648)
  bool cached = false;
  double alpha = 0;
  std::string id;
};

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace

struct pm_context {
  int device = 0;
  int lds_max = 65536;             // LDS bytes a workgroup may use on this device
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;   // PME, correction and bonded work, beside the pair work
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::string err;
  int precision = PM_PRECISION_MIXED;

  // topology
  int n = 0;
  int ntypes = 0;
  std::vector<double> q, mass, ljA, ljB, ljA14, ljB14;
  std::vector<int> type, molecule;
  std::vector<int> ex_a, ex_b, ex_c;       // as given
  std::vector<int> tab_a, tab_b, tab_c;    // ExclusionTable::getTable order
  int max_delta = 0;
  double csf = 1.0;
  bool have_topo = false;

  // boundary
  PmBox box{};
  double cellsize = 0;
  bool have_box = false;
  double realcell[3] = {0, 0, 0};
  int ncells = 0;

  std::vector<ForceDesc> forces;

  // bonded terms
  int nbonds = 0, nangles = 0;
  int *d_b_a = nullptr, *d_b_b = nullptr, *d_b_row = nullptr, *d_b_idx = nullptr;
  double *d_b_k = nullptr, *d_b_r0 = nullptr;
  int *d_a_a = nullptr, *d_a_b = nullptr, *d_a_c = nullptr, *d_a_row = nullptr, *d_a_idx = nullptr;
  double *d_a_k = nullptr, *d_a_t0 = nullptr, *d_a_kub = nullptr, *d_a_rub = nullptr;
  // torsions: kind 0 dihedrals, kind 1 impropers
  int ntors[2] = {0, 0};
  int *d_t_row[2] = {nullptr, nullptr}, *d_t_idx[2] = {nullptr, nullptr}, *d_t_term[2] = {nullptr, nullptr};
  int4* d_t_atoms[2] = {nullptr, nullptr};
  double *d_t_k[2] = {nullptr, nullptr}, *d_t_phase[2] = {nullptr, nullptr};
  int* d_t_n[2] = {nullptr, nullptr};
  std::vector<int> h_ba, h_bb, h_aa, h_ab, h_ac;   // host copies for the constraint list
  std::vector<double> h_br0, h_at0;

  // SHAKE / RATTLE (pm_md_set_constraints, pm_shake, pm_rattle)
  bool cons_built = false;
  int cons_n = 0, cons_ngroup = 0, cons_err_cap = 0;
  int *d_cons_gstart = nullptr, *d_cons_a1 = nullptr, *d_cons_a2 = nullptr, *d_cons_sweeps = nullptr;
  int* d_cons_la = nullptr;
  int4* d_cons_gatom = nullptr;
  double *d_cons_r0 = nullptr, *d_rmass = nullptr, *d_cons_err = nullptr;
  double *d_cons_scratch = nullptr, *d_shake_last = nullptr;
  double *d_cons_x = nullptr, *d_cons_v = nullptr, *d_cons_xl = nullptr;   // pm_shake / pm_rattle staging
  double shake_eps = 0.0, rattle_eps = 0.0;
  int shake_maxit = 0, rattle_maxit = 0;

  // device topology
  double *d_q = nullptr, *d_mass = nullptr;
  int *d_type = nullptr, *d_mol = nullptr;
  int *d_ex_row = nullptr, *d_ex_partner = nullptr, *d_ex_cls = nullptr;      // lookup CSR
  int *d_xs_row = nullptr, *d_xs_part = nullptr;   // resolved exclusions, both directions
  int xs_maxdeg = 0;
  int *d_cx_row = nullptr, *d_cx_partner = nullptr, *d_cx_cls = nullptr;      // correction CSR
  int *d_tab_a = nullptr, *d_tab_b = nullptr, *d_tab_c = nullptr;
  float4* d_lj_f = nullptr;
  double4* d_lj_d = nullptr;
  // per-atom LJ (R = sigma 2^(-5/6), e = sqrt(eps), R14, e14), when set
  float4* d_atom_lj_f = nullptr;
  double4* d_atom_lj_d = nullptr;
  std::vector<double> lj_R, lj_e;

  // adjoint work buffers (Hessian-vector products, parameter gradients,
  // Lanczos), kept across calls and grown on demand
  double4* d_adj_sv = nullptr;          // [adj_nslot] direction or charge per slot
  long long* d_adj_slot = nullptr;      // [3 adj_nslot] per-slot fixed point
  long long* d_adj_fix = nullptr;       // [3 adj_n] per-atom fixed point
  double* d_adj_v = nullptr;            // [3 adj_n] direction
  double* d_adj_out = nullptr;          // [3 adj_n] FP64 parts
  long long adj_nslot = 0;
  int adj_n = 0;

  // state
  double *d_pos = nullptr, *d_vel = nullptr;
  long long* d_force = nullptr;
  double* d_energy = nullptr;     // [PM_NE]
  double* d_scratch3 = nullptr;   // [3n] double conversions

  // cells and slots
  int* d_bin_count = nullptr;        // per-cell atom counts of the binning in progress
  int* d_scan_aux = nullptr;       // [2 per 8192 cells] block totals of the multi-block cell scan
  int *d_cell_of = nullptr, *d_cell_count = nullptr, *d_cell_start = nullptr, *d_cell_slot = nullptr,
      *d_cell_rank = nullptr, *d_cell_atoms = nullptr, *d_cluster_cell = nullptr, *d_counts = nullptr;
  float* d_key = nullptr;
  int4* d_slot_pos = nullptr;
  longlong4* d_slot_pos64 = nullptr;
  int *d_slot_meta = nullptr, *d_atom_of_slot = nullptr;
  long long* d_force_slot = nullptr;   // [3 * (nslot_max + 1)] pair-kernel forces per slot
  int ncl_max = 0, nslot_max = 0;
  int cells_alloc = 0;
  int cells_nz = 0;       // cells per column the slot buffers were sized for
  int m = 8;              // atoms per cluster of the current layout
  int cells_m = 0;        // cluster size the slot buffers were sized for
  int m_override = 0;     // 4 or 8 to force the cluster size (0: the default, 8)

  // stencils keyed by algorithm cutoff
  std::map<double, std::pair<int4*, int>> stencils;
  std::map<double, std::pair<int4*, int>> half_stencils;

  // lists
  int *d_jlist = nullptr, *d_jcount = nullptr, *d_overflow = nullptr;
  unsigned int* d_jmask = nullptr;     // per list entry: FULL / MODIFIED masks of flagged entries
  long long jlist_alloc = 0;
  int lcap = 0;
  unsigned long long* d_stats = nullptr;

  // energy partial records
  double* d_epart = nullptr;
  long long epart_alloc = 0;
  double* d_ered = nullptr;       // reduction stage-1 partials

  // PME buffers (one grid size at a time)
  int pme_nx = 0, pme_ny = 0, pme_nz = 0, pme_prec = -1;
  unsigned long long* d_grid_fix = nullptr;
  void* d_grid_r = nullptr;
  void* d_grid_c = nullptr;
  void* d_theta = nullptr;   // theta(m) on the line-major half grid (grid precision)
  double* d_lat = nullptr;   // lattice vectors of the PME real term in small boxes
  int nlat = 0;
  double lat_key[4] = {0, 0, 0, 0};   // rc and box edges d_lat was built for
  double *d_bmod = nullptr, *d_expt = nullptr, *d_tw = nullptr;
  PmFftPlan plans[3];
  double pme_alpha_tables = -1;

  // MD
  std::vector<int> md_handles;
  int md_integrator = PM_INTEGRATOR_LEAPFROG;
  double md_dt = 1.0, md_temp = 300.0, md_gamma = 0.0;
  unsigned long long md_seed = 0, md_counter = 0;
  bool md_forces_valid = false;
  PmMdDevice* d_md = nullptr;        // noise counter and bath temperature read by the kernels

  // captured MD steps (hipGraph), replayed while nothing they reference changes
  int graph_option = 1;              // PM_OPT_GRAPH
  int pme_fft_option = 0;            // PM_OPT_PME_FFT
  int pme_spread_option = 1;         // PM_OPT_PME_SPREAD (1: fused with the z transform where it fits)
  int pair_lanes_option = 1;         // PM_OPT_PAIR_LANES (1: one lane per list entry)
  int pme_prio_option = -1;          // PM_OPT_PME_PRIORITY (-1: automatic)

  // pair-list reuse (PM_OPT_LIST_SKIN): lists hold j atoms within the pair
  // cutoff + skin and are rebuilt, with the cell layout, when an atom has
  // moved more than skin / 2 since the build (checked on the device)
  int skin_milli = -1;               // -1 automatic, 0 off, else 1e-3 A
  bool disp_prechecked = false;      // the next evaluation's displacement test already queued
  bool md_prebinned = false;         // the next evaluation's binning and force clearing already queued
  int prebin_gen = -1;               // layout_gen the queued binning was made for
  int layout_gen = 0;                // +1 whenever the cell and slot buffers are reallocated
  int* d_rebuild = nullptr;          // device flags: [0] = rebuild in this evaluation, [1] = prune the inner lists
  int4* d_xbuild = nullptr;          // [n] fixed-point position at the last build, w = 1 when valid
  // inner (pruned) pair lists (PM_OPT_PRUNE_SKIN): entries within the pair
  // cutoff + the inner skin, rewritten by the pair kernel whenever an atom has
  // moved more than half the inner skin since the last prune (or on builds)
  int prune_milli = -1;              // -1 automatic, 0 off, else 1e-3 A
  int list_test_option = -1;         // PM_OPT_LIST_TEST: -1 automatic, 0 bounding box where it holds, 1 exact
  int4* d_xprune = nullptr;          // [n] fixed-point position at the last prune, w = 1 when valid
  int* d_ilist = nullptr;
  unsigned int* d_imask = nullptr;
  int* d_icount = nullptr;
  long long ilist_alloc = 0;
  bool inner_last = false;           // the last evaluation's pair pass kept inner lists
  bool list_stale = true;            // x_build must be invalidated before the next reuse
  unsigned long long gen = 1;        // bumped by every allocation and every setter
  hipGraphExec_t md_exec = nullptr;
  int md_exec_steps = 0, md_exec_kind = -1;
  unsigned long long md_exec_gen = 0;

  long long last_stats[6] = {0, 0, 0, 0, 0, 0};

  // kernel timing (HIP events on the context's stream), per category
  unsigned timing_mask = 0;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_open;
  std::vector<hipEvent_t> ev_pool;
  double t_ms[PM_TIMING_CATEGORIES] = {0};
  long long t_count[PM_TIMING_CATEGORIES] = {0};
};

namespace {

void drop_graph(pm_context* c);

int fail(pm_context* c, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHECK(ctx, call)                                                                      \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return fail(ctx, PM_ERR_DEVICE, "HIP error %s at %s:%d (%s)", hipGetErrorString(e_),       \
                  __FILE__, __LINE__, #call);                                                    \
  } while (0)

template <typename T>
int dalloc(pm_context* c, T** p, size_t count) {
  c->gen++;   // device pointers change: captured MD steps are stale
  if (*p) { hipFree(*p); *p = nullptr; }
  if (count == 0) count = 1;
  HIPCHECK(c, hipMalloc((void**)p, count * sizeof(T)));
  return PM_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) hipFree((void*)p);
  p = nullptr;
}

template <typename T>
int upload(pm_context* c, T** p, const std::vector<T>& v) {
  int r = dalloc(c, p, v.size());
  if (r) return r;
  if (!v.empty()) HIPCHECK(c, hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return PM_OK;
}

std::string fmtnum(double v) {
  std::ostringstream os;
  os << v;
  return os.str();
}

// ---------------------------------------------------------------- timing
hipEvent_t take_event(pm_context* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

void harvest_timing(pm_context* c) {
  for (auto& p : c->ev_open) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
      c->t_ms[p.first] += ms;
      c->t_count[p.first] += 1;
    }
    c->ev_pool.push_back(p.second.first);
    c->ev_pool.push_back(p.second.second);
  }
  c->ev_open.clear();
}

struct TimedScope {
  pm_context* c;
  int cat;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  TimedScope(pm_context* ctx, int category, hipStream_t stream = nullptr)
      : c(ctx), cat(category), st(stream ? stream : ctx->stream) {
    if (c->timing_mask & (1u << cat)) {
      if (c->ev_open.size() > 8192) {
        hipStreamSynchronize(c->stream);
        if (c->stream2) hipStreamSynchronize(c->stream2);
        harvest_timing(c);
      }
      a = take_event(c);
      b = take_event(c);
      hipEventRecord(a, st);
    }
  }
  ~TimedScope() {
    if (a) {
      hipEventRecord(b, st);
      c->ev_open.push_back(std::make_pair(cat, std::make_pair(a, b)));
    }
  }
};

// ----------------------------------------------------------- cell helpers
// ArrayCellListStructure::basisCell (topology/ArrayCellListStructure.h:117-122)
int basis(int d, int n) {
  int maxn = kMaxInt2 - kMaxInt2 % n;
  int n1 = -n / 2;
  int maxn1 = maxn - n1;
  return (d + maxn1) % n + n1;
}

// CellListEnumerator<PBC,Cubic>::initialize delta list
// (topology/CellListEnumerator_periodicBoundaries.h:37-108): every cell delta
// whose minimum gap is below the cutoff, mapped to its basis cell,
// deduplicated, zero excluded, sorted lexicographically.
std::vector<int4> delta_list(const pm_context* c, double cutoff) {
  const double cutoff2 = cutoff * cutoff;
  const double* rc = c->realcell;
  const int* n = c->box.n;
  int nx = (int)(cutoff / rc[0] + 1.0 + kEpsilon);
  int ny = (int)(cutoff / rc[1] + 1.0 + kEpsilon);
  int nz = (int)(cutoff / rc[2] + 1.0 + kEpsilon);
  int n0 = std::min(nx, n[0]), n1 = std::min(ny, n[1]), n2 = std::min(nz, n[2]);
  double xx = rc[0] * rc[0], yy = rc[1] * rc[1], zz = rc[2] * rc[2];
  std::set<std::tuple<int, int, int>> seen;
  seen.insert(std::make_tuple(0, 0, 0));
  std::vector<std::tuple<int, int, int>> out;
  for (int k = -n0; k <= n0; k++) {
    int x = std::max(std::abs(k) - 1, 0);
    double d0 = x * x * xx;
    for (int l = -n1; l <= n1; l++) {
      int y = std::max(std::abs(l) - 1, 0);
      double d1 = d0 + y * y * yy;
      for (int m = -n2; m <= n2; m++) {
        int z = std::max(std::abs(m) - 1, 0);
        if (d1 + z * z * zz < cutoff2) {
          auto b = std::make_tuple(basis(k, n[0]), basis(l, n[1]), basis(m, n[2]));
          if (!seen.count(b)) { seen.insert(b); out.push_back(b); }
        }
      }
    }
  }
  std::sort(out.begin(), out.end());
  std::vector<int4> r;
  for (auto& t : out) r.push_back(make_int4(std::get<0>(t), std::get<1>(t), std::get<2>(t), 0));
  return r;
}

// column stencil of the pair lists: the distinct (x, y) offsets, in basis
// form, of the cells within the cutoff of a cell (self, the deltas and their
// negations): the cell columns that can hold an atom within the cutoff of
// an atom of the middle column
std::vector<int4> column_stencil(const pm_context* c, double cutoff) {
  std::vector<int4> h = delta_list(c, cutoff);
  std::set<std::pair<int, int>> seen;
  std::vector<int4> r;
  auto add = [&](int a, int b) {
    auto t = std::make_pair(basis(a, c->box.n[0]), basis(b, c->box.n[1]));
    if (seen.count(t)) return;
    seen.insert(t);
    r.push_back(make_int4(t.first, t.second, 0, 0));
  };
  add(0, 0);
  for (auto& d : h) { add(d.x, d.y); add(-d.x, -d.y); }
  return r;
}

int get_stencil(pm_context* c, double cutoff, bool half, int4** dptr, int* count) {
  auto& m = half ? c->half_stencils : c->stencils;
  auto it = m.find(cutoff);
  if (it == m.end()) {
    std::vector<int4> v = half ? delta_list(c, cutoff) : column_stencil(c, cutoff);
    int4* d = nullptr;
    int r = dalloc(c, &d, v.size());
    if (r) return r;
    if (!v.empty()) HIPCHECK(c, hipMemcpy(d, v.data(), v.size() * sizeof(int4), hipMemcpyHostToDevice));
    it = m.emplace(cutoff, std::make_pair(d, (int)v.size())).first;
  }
  *dptr = it->second.first;
  *count = it->second.second;
  return PM_OK;
}

void free_stencils(pm_context* c) {
  for (auto& kv : c->stencils) hipFree(kv.second.first);
  for (auto& kv : c->half_stencils) hipFree(kv.second.first);
  c->stencils.clear();
  c->half_stencils.clear();
}

// ------------------------------------------------------------ allocation
// atoms per cluster: 8 unless set to 4 (4 shrinks the clusters' bounding
// boxes, so fewer list entries miss the cutoff, at the price of twice the
// j-force atomics per lane; 16 issued 29 % more pair lanes than 8 on the 98k
// water box with 5 A cells and ran 25 % slower, and is no longer built)
int choose_m(const pm_context* c) {
  // per-atom LJ parameters run in the pair kernel's 8-atom-cluster form only
  if (c->d_atom_lj_f || c->d_atom_lj_d) return 8;
  if (c->m_override == 4 || c->m_override == 8) return c->m_override;
  return 8;
}

// the exact per-atom list test in every box (else the bounding-box test where
// the box is wide enough): the option, or the automatic choice (exact: on the
// 98k water box it lists 15 % fewer entries than the bounding box for a build
// 10 % longer, 318.5 against 312.3 ns/day at a 1.4 A skin)
bool list_exact(const pm_context* c) { return c->list_test_option != 0; }

// wave priority of the PME kernels: the option, or the automatic choice
int pme_priority(const pm_context* c) { return c->pme_prio_option >= 0 ? c->pme_prio_option : 2; }

// inner-list skin in A (0: no inner lists): the option, or the automatic
// choice; used only below the list skin
double prune_skin(const pm_context* c) {
  if (c->prune_milli >= 0) return 1e-3 * c->prune_milli;
  return PM_PRUNE_SKIN_AUTO;
}

// pair kernel layout in effect: the lane-per-entry kernel (1) unless the
// option asks for rows (0) or its LDS (type table + four waves' areas) would
// not fit a workgroup (FP64 with over 41 atom types)
static int pair_lanes(const pm_context* c) {
  const size_t tab = (size_t)(c->ntypes + 1) * (c->ntypes + 1) * (c->precision == PM_PRECISION_DOUBLE ? 32 : 16);
  return c->pair_lanes_option && (tab + 15) / 16 * 16 + 4 * PM_JL_LDS <= (size_t)c->lds_max ? 1 : 0;
}

// list skin in A: the option, or the automatic choice
double list_skin(const pm_context* c) {
  if (c->skin_milli >= 0) return 1e-3 * c->skin_milli;
  return c->precision == PM_PRECISION_DOUBLE ? PM_LIST_SKIN_AUTO_DOUBLE : PM_LIST_SKIN_AUTO_MIXED;
}

int ensure_cells(pm_context* c) {
  c->m = choose_m(c);
  if (c->cells_alloc == c->ncells && c->cells_nz == c->box.n[2] && c->cells_m == c->m && c->d_cell_count &&
      c->d_slot_pos)
    return PM_OK;
  int n = c->n, nc = c->ncells;
  int r;
  if (c->box.n[2] > 8192)
    return fail(c, PM_ERR_UNSUPPORTED, "%d cells along z (at most 8192): use a larger cellsize", c->box.n[2]);
  if ((r = dalloc(c, &c->d_cell_of, n))) return r;
  if ((r = dalloc(c, &c->d_key, n))) return r;
  if ((r = dalloc(c, &c->d_cell_count, nc))) return r;
  if ((r = dalloc(c, &c->d_scan_aux, 2 * ((size_t)nc / 4096 + 2)))) return r;
  if ((r = dalloc(c, &c->d_bin_count, nc))) return r;
  HIPCHECK(c, hipMemset(c->d_bin_count, 0, (size_t)nc * sizeof(int)));
  if ((r = dalloc(c, &c->d_cell_start, nc))) return r;
  if ((r = dalloc(c, &c->d_cell_slot, nc))) return r;
  if ((r = dalloc(c, &c->d_cell_rank, n))) return r;
  if ((r = dalloc(c, &c->d_cell_atoms, n))) return r;
  if ((r = dalloc(c, &c->d_counts, 4))) return r;
  const int m = c->m;
  // padding: fewer than m slots after each cell column
  long long slots = (long long)n + (long long)(m - 1) * (nc / c->box.n[2]);
  c->ncl_max = (int)((slots + m - 1) / m);
  c->nslot_max = c->ncl_max * m;
  if ((r = dalloc(c, &c->d_cluster_cell, c->ncl_max))) return r;
  if ((r = dalloc(c, &c->d_slot_pos, (size_t)c->nslot_max + 1))) return r;
  if ((r = dalloc(c, &c->d_slot_pos64, (size_t)c->nslot_max + 1))) return r;
  if ((r = dalloc(c, &c->d_slot_meta, (size_t)c->nslot_max + 1))) return r;
  if ((r = dalloc(c, &c->d_atom_of_slot, (size_t)c->nslot_max + 1))) return r;
  if ((r = dalloc(c, &c->d_force_slot, 3 * ((size_t)c->nslot_max + 1)))) return r;
  if ((r = dalloc(c, &c->d_xbuild, (size_t)n))) return r;
  if ((r = dalloc(c, &c->d_xprune, (size_t)n))) return r;
  if (!c->d_rebuild && (r = dalloc(c, &c->d_rebuild, 2))) return r;
  HIPCHECK(c, hipMemset(c->d_rebuild, 0, 2 * sizeof(int)));
  c->list_stale = true;
  HIPCHECK(c, hipMemset(c->d_counts, 0, 4 * sizeof(int)));
  if (!c->d_stats) {
    if ((r = dalloc(c, &c->d_stats, 4))) return r;
    HIPCHECK(c, hipMemset(c->d_stats, 0, 4 * sizeof(unsigned long long)));
  }
  c->cells_alloc = nc;
  c->cells_nz = c->box.n[2];
  c->cells_m = m;
  c->jlist_alloc = 0;
  c->ilist_alloc = 0;
  c->layout_gen++;
  return PM_OK;
}

// list capacity per cluster: atoms within the list cutoff of a cell-sized box
int list_capacity(const pm_context* c, double cut) {
  double V = c->box.L[0] * c->box.L[1] * c->box.L[2];
  double rho = c->n / V;
  double a = std::max(c->realcell[0], std::max(c->realcell[1], c->realcell[2]));
  double ext = a + 2.0 * cut;
  double est = rho * ext * ext * ext;
  long long cap = (long long)(1.5 * est) + 256;
  cap = std::min(cap, (long long)c->n + 64);
  cap = (cap + 63) / 64 * 64;
  return (int)std::max(cap, 64LL);
}

int ensure_lists(pm_context* c, int lcap) {
  long long need = (long long)c->ncl_max * lcap;
  if (need <= c->jlist_alloc && c->d_jlist) { c->lcap = lcap; return PM_OK; }
  int r;
  if ((r = dalloc(c, &c->d_jlist, (size_t)need))) return r;
  if ((r = dalloc(c, &c->d_jmask, (size_t)need))) return r;
  if ((r = dalloc(c, &c->d_jcount, c->ncl_max))) return r;
  c->jlist_alloc = need;
  c->lcap = lcap;
  return PM_OK;
}

// the inner lists, as large as the full ones (they hold a subset)
int ensure_inner_lists(pm_context* c) {
  const long long need = (long long)c->ncl_max * c->lcap;
  if (need <= c->ilist_alloc && c->d_ilist) return PM_OK;
  int r;
  if ((r = dalloc(c, &c->d_ilist, (size_t)need))) return r;
  if ((r = dalloc(c, &c->d_imask, (size_t)need))) return r;
  if ((r = dalloc(c, &c->d_icount, c->ncl_max))) return r;
  c->ilist_alloc = need;
  return PM_OK;
}

int ensure_epart(pm_context* c, long long records) {
  if (records <= c->epart_alloc && c->d_epart) return PM_OK;
  int r = dalloc(c, &c->d_epart, (size_t)records * PM_NE);
  if (r) return r;
  c->epart_alloc = records;
  return PM_OK;
}

// ---------------------------------------------------------------- PME set-up
// erfc bisection of NonbondedPMEwaldSystemForce::initialize (:564-586)
double pme_alpha(double rc, double accuracy) {
  double alpha = 1.0;
  while (std::erfc(alpha * rc) / rc >= accuracy) alpha *= 2.0;
  double low = 0.0, high = alpha;
  for (unsigned int i = 0; i < 100; ++i) {
    alpha = 0.5 * (low + high);
    if (std::erfc(alpha * rc) / rc >= accuracy) low = alpha; else high = alpha;
  }
  return alpha;
}

// BSpline::set at w = 0 (base/BSpline.cpp:64-117)
void bspline_host(int order, double w, double* th) {
  std::vector<double> t(order), d(order);
  if (order == 2) { th[1] = w; th[0] = 1 - w; return; }
  t[order - 1] = 0.0; t[1] = w; t[0] = 1.0 - w;
  for (int k = 2; k < order - 1; k++) {
    double div = 1.0 / k;
    t[k] = div * w * t[k - 1];
    for (int j = 1; j < k; j++) t[k - j] = div * ((w + j) * t[k - j - 1] + (k + 1 - j - w) * t[k - j]);
    t[0] = div * (1 - w) * t[0];
  }
  double div = 1.0 / (order - 1.0);
  t[order - 1] = div * w * t[order - 2];
  for (int j = 1; j < order - 1; j++)
    t[order - 1 - j] = div * ((w + j) * t[order - j - 2] + (order - j - w) * t[order - 1 - j]);
  t[0] = div * (1 - w) * t[0];
  for (int k = 0; k < order; k++) th[k] = t[k];
}

// Grid::dftmod (base/Grid.h:650-665)
void dftmod(int order, int n, const double* theta, double* mod) {
  for (int i = 0; i < n; i++) {
    double sc = 0.0, ss = 0.0;
    for (int j = 0; j < order; j++) {
      double x = kPi * 2.0 * i * j / (double)n;
      sc += theta[j] * std::cos(x);
      ss += theta[j] * std::sin(x);
    }
    mod[i] = 1.0 / (sc * sc + ss * ss);
  }
}

// the fused kz-plane transform (PM_OPT_PME_FFT 2) where a plane fits in LDS.
// The automatic choice is the separate passes: the plane kernel has one
// block per kz plane (49 for a 96^3 grid), and beside the pair kernel on the
// other stream those few blocks take several times longer than the passes'
// hundreds.
int pme_plane(const pm_context* c, const ForceDesc& f) {
  if (c->pme_fft_option != 2) return 0;
  return pm_pme_plane_fits(f.nx, f.ny, c->precision, c->lds_max);
}

bool fft_plan(int n, PmFftPlan& p) {
  p.n = n;
  p.nst = 0;
  int m = n;
  while (m % 4 == 0 && m != 8) { p.rad[p.nst++] = 4; m /= 4; if (p.nst >= PM_FFT_MAXST) return false; }
  while (m % 2 == 0) { p.rad[p.nst++] = 2; m /= 2; if (p.nst >= PM_FFT_MAXST) return false; }
  while (m % 3 == 0) { p.rad[p.nst++] = 3; m /= 3; if (p.nst >= PM_FFT_MAXST) return false; }
  while (m % 5 == 0) { p.rad[p.nst++] = 5; m /= 5; if (p.nst >= PM_FFT_MAXST) return false; }
  while (m % 7 == 0) { p.rad[p.nst++] = 7; m /= 7; if (p.nst >= PM_FFT_MAXST) return false; }
  return m == 1;
}

int ensure_pme(pm_context* c, ForceDesc& f) {
  if (!f.cached) {
    f.alpha = f.alpha_in <= 0.0 ? pme_alpha(f.rc, f.accuracy) : f.alpha_in;
    f.cached = true;
  }
  const int nx = f.nx, ny = f.ny, nz = f.nz, nzc = nz / 2 + 1;
  const bool same_grid = c->pme_nx == nx && c->pme_ny == ny && c->pme_nz == nz && c->pme_prec == c->precision;
  if (!f.recip) return PM_OK;
  if (!same_grid) {
    if (nx > 4096 || ny > 4096 || nz > 4096 || nx < 8 || ny < 8 || nz < 8)
      return fail(c, PM_ERR_UNSUPPORTED, "PME grid %d x %d x %d: 8 to 4096 points per axis", nx, ny, nz);
    if (!fft_plan(nx, c->plans[0]) || !fft_plan(ny, c->plans[1]) || !fft_plan(nz, c->plans[2]))
      return fail(c, PM_ERR_UNSUPPORTED, "PME grid %d x %d x %d: the GPU FFT handles lengths with factors 2, 3, 5 and 7 only", nx, ny, nz);
    size_t ng = (size_t)nx * ny * nz, ngc = (size_t)nx * ny * nzc;
    size_t er = c->precision == PM_PRECISION_DOUBLE ? sizeof(double) : sizeof(float);
    int r;
    if ((r = dalloc(c, &c->d_grid_fix, ng))) return r;
    HIPCHECK(c, hipMemset(c->d_grid_fix, 0, ng * sizeof(unsigned long long)));
    if (c->d_grid_r) hipFree(c->d_grid_r);
    if (c->d_grid_c) hipFree(c->d_grid_c);
    if (c->d_theta) hipFree(c->d_theta);
    c->d_grid_r = c->d_grid_c = c->d_theta = nullptr;
    HIPCHECK(c, hipMalloc(&c->d_grid_r, ng * er));
    HIPCHECK(c, hipMalloc(&c->d_grid_c, ngc * er * 2));
    HIPCHECK(c, hipMalloc(&c->d_theta, ngc * er));
    // twiddles exp(-2 pi i m / n) for the three lengths
    std::vector<double> tw;
    int off = 0;
    for (int a = 0; a < 3; a++) {
      int len = c->plans[a].n;
      c->plans[a].tw = off;
      for (int m = 0; m < len; m++) {
        double ang = -2.0 * kPi * (double)m / len;
        tw.push_back(std::cos(ang));
        tw.push_back(std::sin(ang));
      }
      off += len;
    }
    if ((r = upload(c, &c->d_tw, tw))) return r;
    c->pme_nx = nx; c->pme_ny = ny; c->pme_nz = nz; c->pme_prec = c->precision;
    c->pme_alpha_tables = -1;
  }
  if (c->pme_alpha_tables != f.alpha) {
    // Grid::initialize (base/Grid.h:557-648)
    const double lx = c->box.L[0], ly = c->box.L[1], lz = c->box.L[2];
    const double fac = kPi * kPi / (f.alpha * f.alpha);
    std::vector<double> bm(nx + ny + nz), ex(nx + ny + nzc);
    std::vector<double> th(f.order);
    bspline_host(f.order, 0.0, th.data());
    dftmod(f.order, nx, th.data(), bm.data());
    dftmod(f.order, ny, th.data(), bm.data() + nx);
    dftmod(f.order, nz, th.data(), bm.data() + nx + ny);
    const double wr = 1.0 / lx, lr = 1.0 / ly, hr = 1.0 / lz;
    for (int i = 0; i < nx; i++) { int i0 = i <= nx / 2 ? i : i - nx; double t = i0 * wr; ex[i] = std::exp(-fac * (t * t)); }
    for (int j = 0; j < ny; j++) { int j0 = j <= ny / 2 ? j : j - ny; double t = j0 * lr; ex[nx + j] = std::exp(-fac * (t * t)); }
    for (int k = 0; k < nzc; k++) { double t = k * hr; ex[nx + ny + k] = std::exp(-fac * (t * t)); }
    int r;
    if ((r = upload(c, &c->d_bmod, bm))) return r;
    if ((r = upload(c, &c->d_expt, ex))) return r;
    // theta(m) for the x passes that write no energy records
    PmPmeLaunch t{};
    t.nx = nx; t.ny = ny; t.nz = nz; t.nzc = nzc;
    for (int k = 0; k < 3; k++) t.Lr[k] = 1.0 / c->box.L[k];
    t.bmod = c->d_bmod;
    t.expf_tab = c->d_expt;
    t.inv_pi_v = 1.0 / (kPi * (c->box.L[0] * c->box.L[1] * c->box.L[2]));
    t.theta = c->d_theta;
    pm_launch_pme_theta(c->stream, t, c->precision);
    HIPCHECK(c, hipGetLastError());
    c->pme_alpha_tables = f.alpha;
  }
  return PM_OK;
}

// ------------------------------------------------------------ force parsing
std::vector<std::string> tokens(const std::string& s) {
  std::istringstream is(s);
  std::vector<std::string> t;
  std::string w;
  while (is >> w) t.push_back(w);
  return t;
}

bool ieq(const std::string& a, const std::string& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); i++)
    if (tolower((unsigned char)a[i]) != tolower((unsigned char)b[i])) return false;
  return true;
}

int parse_force(pm_context* c, const char* id_c, const double* v, int nv, ForceDesc& f) {
  std::string id(id_c ? id_c : "");
  std::vector<std::string> t = tokens(id);
  if (t.size() == 1 && (ieq(t[0], "Bond") || ieq(t[0], "Angle") || ieq(t[0], "Dihedral") || ieq(t[0], "Improper"))) {
    if (nv != 0) return fail(c, PM_ERR_ARGUMENT, "force '%s' takes no parameters", id.c_str());
    f.bonded = ieq(t[0], "Bond") ? 1 : ieq(t[0], "Angle") ? 2 : ieq(t[0], "Dihedral") ? 3 : 4;
    f.id = id;
    return PM_OK;
  }
  if (t.size() == 1 && ieq(t[0], "CoulombPME")) {
    // alias of the full PME force (registerForceExemplarsFastElectrostatic.cpp:50)
    t = tokens("Coulomb -algorithm PMEwald -real -reciprocal -correction -interpolation BSpline");
  }
  std::vector<int> pots;
  std::vector<int> swfs;
  std::string algo;
  for (size_t i = 0; i < t.size(); i++) {
    const std::string& w = t[i];
    if (ieq(w, "LennardJones")) pots.push_back(0);
    else if (ieq(w, "Coulomb")) pots.push_back(1);
    else if (ieq(w, "CoulombEwaldReal")) pots.push_back(2);
    else if (ieq(w, "-algorithm") && i + 1 < t.size()) algo = t[++i];
    else if (ieq(w, "-switchingFunction") && i + 1 < t.size()) {
      const std::string& s = t[++i];
      if (ieq(s, "C1")) swfs.push_back(PM_SWF_C1);
      else if (ieq(s, "C2")) swfs.push_back(PM_SWF_C2);
      else if (ieq(s, "Cutoff")) swfs.push_back(PM_SWF_CUTOFF);
      else return fail(c, PM_ERR_UNSUPPORTED, "Force '%s': switching function %s is not implemented on the GPU", id.c_str(), s.c_str());
    } else if (ieq(w, "-real")) f.real = true;
    else if (ieq(w, "-reciprocal")) f.recip = true;
    else if (ieq(w, "-correction")) f.corr = true;
    else if (ieq(w, "-interpolation") && i + 1 < t.size()) {
      if (!ieq(t[++i], "BSpline"))
        return fail(c, PM_ERR_UNSUPPORTED, "Force '%s': only BSpline interpolation is implemented on the GPU", id.c_str());
    } else {
      return fail(c, PM_ERR_UNSUPPORTED, "Could not find any match for '%s' in ForceFactory", id.c_str());
    }
  }
  f.id = id;
  if (ieq(algo, "PMEwald")) {
    // NonbondedPMEwaldSystemForce::doMake (:1100-1139), periodic: 7 values
    if (pots.size() != 1 || pots[0] != 1 || !swfs.empty())
      return fail(c, PM_ERR_UNSUPPORTED, "Could not find any match for '%s' in ForceFactory", id.c_str());
    if (!f.real && !f.recip && !f.corr)
      return fail(c, PM_ERR_UNSUPPORTED, "Could not find any match for '%s' in ForceFactory", id.c_str());
    if (f.real && !f.recip && f.corr)
      ;  // registered (true,false,true)
    if (nv != 7) return fail(c, PM_ERR_ARGUMENT, "force PMEwald : expected 7 parameter values, got %d.", nv);
    f.pme = true;
    f.nx = (int)v[0]; f.ny = (int)v[1]; f.nz = (int)v[2];
    f.rc = v[3]; f.order = (int)v[4]; f.accuracy = v[5]; f.alpha_in = v[6];
    std::string err;
    const std::string kw = "PMEwald";
    if (f.order < 2 || (f.order % 2) != 0) err += kw + " order (=" + fmtnum(v[4]) + ") > 1 and even.";
    if (f.nx < f.order || f.ny < f.order || f.nz < f.order)
      err += kw + " force: " + fmtnum(v[4]) + " <= nx (=" + fmtnum(v[0]) + "), " + fmtnum(v[4]) + " <= ny (=" +
             fmtnum(v[1]) + "), " + fmtnum(v[4]) + " <= nz (=" + fmtnum(v[2]) + ").";
    if (f.rc <= 0.0) err += kw + " cutoff (=" + fmtnum(v[3]) + ") > 0.";
    if (f.accuracy <= 0.0) err += kw + " accuracy (=" + fmtnum(v[5]) + ") > 0.";
    if (!err.empty()) return fail(c, PM_ERR_ARGUMENT, " force %s :%s", kw.c_str(), err.c_str());
    if (f.order > 12)
      return fail(c, PM_ERR_UNSUPPORTED, "PME order %d: the GPU kernels implement the even orders 2 to 12", f.order);
    return PM_OK;
  }
  if (!ieq(algo, "NonbondedCutoff"))
    return fail(c, PM_ERR_UNSUPPORTED, "Could not find any match for '%s' in ForceFactory", id.c_str());
  if (pots.empty() || pots.size() > 2 || swfs.size() != pots.size())
    return fail(c, PM_ERR_UNSUPPORTED, "Could not find any match for '%s' in ForceFactory", id.c_str());
  if (pots.size() == 2 && (pots[0] != 0 || pots[1] == 0))
    return fail(c, PM_ERR_UNSUPPORTED, "Could not find any match for '%s' in ForceFactory", id.c_str());
  // only the combinations the periodic factory registers
  // (factories/registerForceExemplarsCutoff.cpp:39,40,45,57,62,63,67,76,88,90,91,131)
  {
    bool ok;
    if (pots.size() == 1)
      ok = pots[0] != 2 || swfs[0] == PM_SWF_CUTOFF;   // CoulombEwaldReal only with Cutoff (:57)
    else if (pots[1] == 2)
      ok = swfs[1] == PM_SWF_CUTOFF;                    // LJ {C1,C2,Cutoff} + EwaldReal/Cutoff (:76,90,131)
    else
      ok = swfs[0] == PM_SWF_C2 && (swfs[1] == PM_SWF_C1 || swfs[1] == PM_SWF_CUTOFF);  // (:88, :91)
    if (!ok) return fail(c, PM_ERR_UNSUPPORTED, "Could not find any match for '%s' in ForceFactory", id.c_str());
  }
  // positional values: per term potential params then switching params, then
  // the algorithm cutoff (OneAtomPair(Two)::getParameters, NonbondedCutoffForce.h:113-117)
  int k = 0;
  for (size_t i = 0; i < pots.size(); i++) {
    Term tm;
    tm.pot = pots[i];
    tm.swf = swfs[i];
    if (tm.pot == 2) {
      if (k >= nv) return fail(c, PM_ERR_ARGUMENT, "force '%s': missing -alpha", id.c_str());
      tm.alpha = v[k++];
      if (!(tm.alpha > 0)) return fail(c, PM_ERR_ARGUMENT, "CoulombEwaldReal -alpha (=%s) > 0.", fmtnum(tm.alpha).c_str());
    }
    if (tm.swf == PM_SWF_C2) {
      if (k + 1 >= nv) return fail(c, PM_ERR_ARGUMENT, "force '%s': missing C2 parameters", id.c_str());
      tm.switchon = v[k++];
      tm.cutoff = v[k++];
      // C2SwitchingFunction::make (topology/C2SwitchingFunction.cpp:32-41)
      if (tm.switchon < 0.0 || tm.cutoff <= 0.0 || tm.switchon >= tm.cutoff)
        return fail(c, PM_ERR_ARGUMENT, "C2 switching function: 0 <= switchon (=%s) < cutoff (=%s).",
                    fmtnum(tm.switchon).c_str(), fmtnum(tm.cutoff).c_str());
    } else {
      if (k >= nv) return fail(c, PM_ERR_ARGUMENT, "force '%s': missing switching cutoff", id.c_str());
      tm.cutoff = v[k++];
      if (!(tm.cutoff > 0))
        return fail(c, PM_ERR_ARGUMENT, "%s swf cutoff (=%s) > 0.", tm.swf == PM_SWF_C1 ? "C1" : "cutoff",
                    fmtnum(tm.cutoff).c_str());
    }
    f.terms.push_back(tm);
  }
  if (k != nv - 1) return fail(c, PM_ERR_ARGUMENT, "force '%s': expected %d parameter values, got %d", id.c_str(), k + 1, nv);
  f.algo_cutoff = v[k];
  if (!(f.algo_cutoff > 0)) return fail(c, PM_ERR_ARGUMENT, "algorithm cutoff (=%s) > 0.", fmtnum(f.algo_cutoff).c_str());
  return PM_OK;
}

void fill_swf(const Term& t, double& cut2, double& on2, double& sw1, double& sw2, double& sw3, double& c15_1,
              double& c05_3, double& c15_3) {
  const double c = t.cutoff, s = t.switchon;
  cut2 = c * c;
  on2 = s * s;
  const double d = c * c - s * s;
  const double d3 = d * d * d;
  sw1 = 1.0 / d3;
  sw2 = c * c - 3.0 * s * s;
  sw3 = 4.0 / d3;
  c15_1 = 1.5 / c;
  c05_3 = 0.5 / (c * c * c);
  c15_3 = 1.5 / (c * c * c);
}

// pair parameters for one kernel pass over the given terms (LJ first)
PmPairParams pair_params(const std::vector<Term>& terms, double csf) {
  PmPairParams p{};
  p.lj_swf = PM_SWF_NONE;
  p.coul_kind = PM_COUL_NONE;
  p.coul_swf = PM_SWF_CUTOFF;
  p.csf = csf;
  p.cut2 = 0;
  bool modify = false;
  for (const Term& t : terms) {
    double cut2, on2, sw1, sw2, sw3, a1, a2, a3;
    fill_swf(t, cut2, on2, sw1, sw2, sw3, a1, a2, a3);
    p.cut2 = std::max(p.cut2, cut2);
    if (t.swf != PM_SWF_CUTOFF) modify = true;
    if (t.pot == 0) {
      p.lj_swf = t.swf;
      p.lj_cut2 = cut2; p.lj_on2 = on2; p.lj_sw1 = sw1; p.lj_sw2 = sw2; p.lj_sw3 = sw3;
      p.lj_c15_1 = a1; p.lj_c05_3 = a2; p.lj_c15_3 = a3;
    } else {
      p.coul_kind = t.pot == 1 ? PM_COUL_PLAIN : PM_COUL_EWALD;
      p.coul_swf = t.swf;
      p.c_cut2 = cut2; p.c_on2 = on2; p.c_sw1 = sw1; p.c_sw2 = sw2; p.c_sw3 = sw3;
      p.c_c15_1 = a1; p.c_c05_3 = a2; p.c_c15_3 = a3;
      if (t.pot == 2) { p.alpha = t.alpha; p.two_alpha_pi = 2.0 * t.alpha / std::sqrt(kPi); }
    }
  }
  p.modify = modify ? 1 : 0;
  return p;
}

struct PairPass {
  std::vector<Term> terms;
  double algo_cutoff;
};

// Plan the pair passes of a force group: one pass per force, except that a
// single-term LennardJones force and a single-term Coulomb force with the same
// algorithm cutoff and the same switching cutoff share one pass (the sum of
// two OneAtomPair evaluations over the same pair set), and that a two-term
// force whose terms switch (OneAtomPairTwo.h:166-178) at different cutoffs
// runs as one pass per term: each switched term is zero beyond its own
// cutoff, so the sum of the two passes is the OneAtomPairTwo evaluation, and
// the pair kernel never clips a term below its pass's cutoff.
// blocks (and energy records) of the lattice real-term kernel
int lattice_blocks(int n) { return std::max(1, std::min(n, 512)); }

// the PME real term needs lattice images beyond the minimal one when a box
// edge is at most twice the cutoff (NonbondedPMEwaldSystemForce.h:630-631)
bool pme_lattice(const pm_context* c, const ForceDesc& f) {
  const double lmin = std::min(c->box.L[0], std::min(c->box.L[1], c->box.L[2]));
  return lmin <= 2.0 * f.rc;
}

// lattice vectors of norm <= 2 rc (PeriodicBoundaryConditions.cpp:80-111),
// (0,0,0) first, then by norm; equal norms keep the generation order (the
// reference's std::sort leaves their order unspecified, which matters only
// for a MODIFIED pair whose primary image lies beyond the cutoff)
int ensure_lattice(pm_context* c, double rc) {
  const double key[4] = {rc, c->box.L[0], c->box.L[1], c->box.L[2]};
  if (c->d_lat && std::equal(key, key + 4, c->lat_key)) return PM_OK;
  const double lx = c->box.L[0], ly = c->box.L[1], lz = c->box.L[2];
  const int bk = (int)std::floor(rc / lx + 1.0), bl = (int)std::floor(rc / ly + 1.0), bm = (int)std::floor(rc / lz + 1.0);
  std::vector<std::array<double, 3>> v;
  for (int k = -bk; k <= bk; k++)
    for (int l = -bl; l <= bl; l++)
      for (int m = -bm; m <= bm; m++) {
        if (k == 0 && l == 0 && m == 0) continue;
        const double x = lx * k, y = ly * l, z = lz * m;
        if (std::sqrt(x * x + y * y + z * z) <= 2 * rc) v.push_back({x, y, z});
      }
  std::stable_sort(v.begin(), v.end(), [](const std::array<double, 3>& a, const std::array<double, 3>& b) {
    return a[0] * a[0] + a[1] * a[1] + a[2] * a[2] < b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
  });
  if (v.size() + 1 > 1024)
    return fail(c, PM_ERR_UNSUPPORTED, "PME real term: %d lattice images within 2 rc (at most 1024; the box is too small "
                                       "for the cutoff)", (int)v.size() + 1);
  std::vector<double> flat = {0.0, 0.0, 0.0};
  for (auto& a : v) flat.insert(flat.end(), a.begin(), a.end());
  int r;
  if ((r = upload(c, &c->d_lat, flat))) return r;
  c->nlat = (int)v.size() + 1;
  std::copy(key, key + 4, c->lat_key);
  return PM_OK;
}

std::vector<PairPass> plan_passes(pm_context* c, const std::vector<int>& handles) {
  std::vector<PairPass> passes;
  std::vector<bool> used(handles.size(), false);
  for (size_t i = 0; i < handles.size(); i++) {
    ForceDesc& f = c->forces[handles[i]];
    if (f.pme || f.bonded || used[i]) continue;
    used[i] = true;
    if (f.terms.size() == 2 && f.terms[0].cutoff != f.terms[1].cutoff &&
        (f.terms[0].swf != PM_SWF_CUTOFF || f.terms[1].swf != PM_SWF_CUTOFF)) {
      // a term with Cutoff switching is still zeroed beyond its cutoff when
      // the other term switches (apply_switch with the Cutoff function)
      passes.push_back(PairPass{{f.terms[0]}, f.algo_cutoff});
      passes.push_back(PairPass{{f.terms[1]}, f.algo_cutoff});
      continue;
    }
    PairPass p{f.terms, f.algo_cutoff};
    if (f.terms.size() == 1) {
      for (size_t j = i + 1; j < handles.size(); j++) {
        ForceDesc& g = c->forces[handles[j]];
        if (used[j] || g.pme || g.bonded || g.terms.size() != 1) continue;
        const Term& a = f.terms[0];
        const Term& b = g.terms[0];
        bool lj_coul = (a.pot == 0) != (b.pot == 0);
        if (lj_coul && g.algo_cutoff == f.algo_cutoff && a.cutoff == b.cutoff) {
          p.terms = a.pot == 0 ? std::vector<Term>{a, b} : std::vector<Term>{b, a};
          used[j] = true;
          break;
        }
      }
    }
    passes.push_back(p);
  }
  // PME real-space terms run as a CoulombEwaldReal / Cutoff pair pass (in
  // boxes no wider than twice the cutoff, k_ewald_lattice instead)
  for (size_t i = 0; i < handles.size(); i++) {
    ForceDesc& f = c->forces[handles[i]];
    if (!f.pme || !f.real || pme_lattice(c, f)) continue;
    Term t;
    t.pot = 2; t.swf = PM_SWF_CUTOFF; t.alpha = f.alpha; t.cutoff = f.rc;
    passes.push_back(PairPass{{t}, f.rc});
  }
  return passes;
}

// ---------------------------------------------------------- evaluation core
// Queues one force-group evaluation at the positions in c->d_pos: forces into
// c->d_force (fixed point, zeroed first), energies into c->d_energy.
// Adjoint evaluation request (pm_hessian_vec, pm_param_gradients): mode 1
// H v, mode 2 parameter gradients; device buffers sized by the caller.
struct AdjointSpec {
  int mode = 0;
  const double* v = nullptr;   // [3n] direction (mode 1)
  double4* slot_v = nullptr;   // [nslot_max + 1]
  long long* slot = nullptr;   // [3 (nslot_max + 1)] per-slot fixed point
  long long* fix = nullptr;    // [3n] per-atom fixed point
  double* hv = nullptr;        // [3n] mode 1 (PME and correction parts, FP64)
  double* dq = nullptr;        // [n] mode 2 (PME, correction and self parts, FP64)
};

// List reuse applies to one pair pass whose algorithm cutoff covers its pair
// cutoff (the reference's cell pairs then hold every pair within the cutoff,
// so the pair set does not depend on the cells); never for adjoint passes.
// Lists hold every j atom within the pair cutoff + skin of some atom of the
// cluster in the minimal image at the build (the bounding-box test in boxes
// wider than 2 (cutoff + skin + cell diagonal), the exact per-atom test in
// smaller ones); a pair within the cutoff in its minimal image later was
// within the cutoff + skin in that image, so in its minimal image, at the
// build while every atom stays within half the skin of its build position.
bool list_reuse(const pm_context* c, const std::vector<PairPass>& passes) {
  const double skin = list_skin(c);
  if (!(skin > 0.0) || passes.size() != 1) return false;
  const double pc = std::sqrt(pair_params(passes[0].terms, c->csf).cut2);
  return passes[0].algo_cutoff >= pc;
}

// Inner lists: with list reuse, an inner skin below the list skin
bool inner_lists(const pm_context* c) {
  const double ps = prune_skin(c);
  return ps > 0.0 && ps < list_skin(c) && c->d_xprune && c->d_rebuild;
}

// The start of the MD forces' next evaluation, run by the position update
// that precedes it (PmStepWork): the binning of the new positions and, when
// the lists are reused and need no invalidation, the displacement test.
// False when the cell layout is not set up yet (the evaluation then does
// both).  On true the update also clears the forces after reading them, and
// the next evaluate_group launches neither the binning nor the clearing (nor
// the test, when it was queued).
bool md_step_work(pm_context* c, PmStepWork* w) {
  *w = PmStepWork{};
  if (!c->have_box || !c->d_bin_count || !c->d_cell_of || !c->d_cell_rank || !c->d_key ||
      c->cells_alloc != c->ncells || c->cells_nz != c->box.n[2] || c->cells_m != choose_m(c))
    return false;
  w->box = c->box;
  w->cell_of = c->d_cell_of;
  w->rank = c->d_cell_rank;
  w->key = c->d_key;
  w->bin_count = c->d_bin_count;
  c->md_prebinned = true;
  c->prebin_gen = c->layout_gen;
  if (!c->list_stale && c->d_xbuild && c->d_rebuild && list_reuse(c, plan_passes(c, c->md_handles))) {
    const double half = 0.5 * list_skin(c);
    for (int k = 0; k < 3; k++) {
      w->to_fix[k] = c->box.to_fix[k];
      w->fix_to_A[k] = c->box.fix_to_A[k];
    }
    w->x_build = c->d_xbuild;
    w->lim2 = (float)half * (float)half;
    w->flag = c->d_rebuild;
    if (inner_lists(c)) {
      const double hin = 0.5 * prune_skin(c);
      w->x_prune = c->d_xprune;
      w->lim2_in = (float)hin * (float)hin;
    }
    c->disp_prechecked = true;
  }
  return true;
}


int evaluate_group(pm_context* c, const std::vector<int>& handles, int flags, bool want_energy = true,
                   const AdjointSpec* adj = nullptr) {
  hipStream_t s = c->stream;
  const int n = c->n;
  const bool vir = (flags & PM_EVAL_VIRIAL) != 0;
  // the position update before this evaluation may have queued the
  // displacement test already (md_disp_check); taken here, before any early
  // return, so that it never carries over to another evaluation
  const bool prechecked = c->disp_prechecked;
  const bool prebinned = c->md_prebinned;
  c->disp_prechecked = false;
  c->md_prebinned = false;
  int r;
  for (int h : handles) {
    ForceDesc& f = c->forces[h];
    if (f.pme) {
      if ((r = ensure_pme(c, f))) return r;
      if (f.real && pme_lattice(c, f) && (r = ensure_lattice(c, f.rc))) return r;
    }
  }
  std::vector<PairPass> passes = plan_passes(c, handles);
  bool any_pme_recip = false;
  for (int h : handles) if (c->forces[h].pme && c->forces[h].recip) any_pme_recip = true;

  // binning and slot layout (shared by every pass and by PME)
  if ((r = ensure_cells(c))) return r;
  if (!c->d_force || !c->d_pos || !c->d_q || !c->d_type || !c->d_stats || !c->d_overflow || !c->d_energy ||
      !c->d_rebuild || !c->d_xbuild)
    return fail(c, PM_ERR_STATE, "internal: device buffer missing before launch");
  const int nc = c->ncells;
  // list reuse: one pair pass whose algorithm cutoff covers its pair cutoff
  // (the reference's cell pairs then hold every pair within the cutoff, so
  // the pair set does not depend on the cells), in a box where the skin
  // keeps the minimal image unique; never for adjoint passes
  const double skin = list_skin(c);
  const bool reuse = !adj && list_reuse(c, passes);
  const int* rb = reuse ? c->d_rebuild : nullptr;
  const bool inner = reuse && inner_lists(c);
  c->inner_last = inner;
  {
  TimedScope ts(c, PM_T_BIN);
  if (reuse && c->list_stale) {
    pm_launch_invalidate_build(s, n, c->d_xbuild);
    pm_launch_invalidate_build(s, n, c->d_xprune);
    c->list_stale = false;
  }
  // (n = 0 without reuse: nothing runs; a recorded replica step keeps its
  // kernel sequence the same with and without reuse)
  const float half = (float)(0.5 * skin);
  const float hin = (float)(0.5 * prune_skin(c));
  pm_launch_check_disp(s, c->box, reuse && !prechecked ? n : 0, c->d_pos, c->d_xbuild, half * half, c->d_rebuild,
                       inner ? c->d_xprune : nullptr, hin * hin);
  // (the MD position update may have binned the atoms and cleared the forces
  // already, for the layout in place; the last fold cleared the slot forces)
  const bool binned = prebinned && c->prebin_gen == c->layout_gen;
  if (!binned) {
    pm_launch_zero_force_cells(s, rb, c->d_force, 3LL * n, c->d_force_slot,
                               passes.empty() ? 0 : 3LL * c->nslot_max, nullptr, 0);
    pm_launch_bin(s, c->box, n, c->d_pos, c->d_cell_of, c->d_bin_count, c->d_cell_rank, c->d_key, rb);
  }
  pm_launch_scan_cells(s, nc, c->box.n[2], c->d_bin_count, c->d_cell_count, c->d_cell_start, c->d_cell_slot, c->d_counts,
                       c->d_cluster_cell, c->ncl_max, c->m, rb, c->d_scan_aux);
  pm_launch_scatter(s, n, c->d_cell_of, c->d_cell_rank, c->d_cell_start, c->d_cell_atoms, rb);
  pm_launch_fill_slots(s, c->box, nc, c->d_pos, c->d_q, c->d_type, c->d_cell_count, c->d_cell_start,
                       c->d_cell_slot, c->d_cell_atoms, c->d_key, c->d_slot_pos,
                       c->precision == PM_PRECISION_DOUBLE ? c->d_slot_pos64 : nullptr, c->d_slot_meta,
                       c->d_atom_of_slot, c->ntypes, c->nslot_max, c->m, rb, reuse ? c->d_xbuild : nullptr,
                       inner ? c->d_xprune : nullptr);
  if (!reuse) c->list_stale = true;   // a layout built without the skin is not reusable
  }

  // energy records: pair passes, PME x-pass blocks, correction blocks
  long long rec = (long long)passes.size() * c->ncl_max;
  std::vector<std::pair<int, long long>> pme_rec;   // (handle, offset)
  for (int h : handles) {
    ForceDesc& f = c->forces[h];
    if (!f.pme) continue;
    long long off = rec;
    if (f.recip) rec += pm_pme_xpass_blocks(f.nx, f.ny, f.nz, c->precision, pme_plane(c, f));
    if (f.corr) rec += ((long long)c->tab_a.size() + 255) / 256;
    if (f.real && pme_lattice(c, f)) rec += lattice_blocks(n);
    pme_rec.push_back(std::make_pair(h, off));
  }
  bool do_bond = false, do_angle = false, do_tors[2] = {false, false};
  for (int h : handles) {
    if (c->forces[h].bonded == 1) do_bond = true;
    if (c->forces[h].bonded == 2) do_angle = true;
    if (c->forces[h].bonded == 3) do_tors[0] = true;
    if (c->forces[h].bonded == 4) do_tors[1] = true;
  }
  if (do_bond && c->nbonds == 0) do_bond = false;
  if (do_angle && c->nangles == 0) do_angle = false;
  for (int k = 0; k < 2; k++)
    if (do_tors[k] && c->ntors[k] == 0) do_tors[k] = false;
  const bool any_bonded = do_bond || do_angle || do_tors[0] || do_tors[1];
  const long long bonded_off = rec;
  if (do_bond) rec += (c->nbonds + 255) / 256;
  if (do_angle) rec += (c->nangles + 255) / 256;
  for (int k = 0; k < 2; k++)
    if (do_tors[k]) rec += (c->ntors[k] + 255) / 256;
  if ((r = ensure_epart(c, std::max(rec, 1LL)))) return r;

  // host-side set-up of every launch first (stencils, list capacity, PME
  // bricks), so that a failure leaves no work half queued
  double lmin_half = 0.5 * std::min(c->box.L[0], std::min(c->box.L[1], c->box.L[2]));
  double cdiag = std::sqrt(c->realcell[0] * c->realcell[0] + c->realcell[1] * c->realcell[1] +
                           c->realcell[2] * c->realcell[2]);
  std::vector<std::pair<PmNbLaunch, PmPairParams>> pair_launch;
  for (size_t pi = 0; pi < passes.size(); pi++) {
    PairPass& ps = passes[pi];
    PmPairParams pp = pair_params(ps.terms, c->csf);
    int4* st;
    int nst;
    const double lskin = reuse ? skin : 0.0;
    if ((r = get_stencil(c, ps.algo_cutoff + lskin, false, &st, &nst))) return r;
    if (nst > 650)
      return fail(c, PM_ERR_UNSUPPORTED, "cell stencil of %d columns (cellsize %g, cutoff %g): use a cellsize of at least cutoff/4", nst, c->cellsize, ps.algo_cutoff);
    double pcut = std::sqrt(pp.cut2);
    if ((r = ensure_lists(c, list_capacity(c, pcut + lskin)))) return r;
    if (inner && (r = ensure_inner_lists(c))) return r;
    pair_launch.push_back(std::make_pair(PmNbLaunch{}, pp));
    PmNbLaunch& a = pair_launch.back().first;
    a.slot_pos = c->d_slot_pos;
    a.slot_pos64 = c->d_slot_pos64;
    a.slot_meta = c->d_slot_meta;
    a.cluster_cell = c->d_cluster_cell;
    a.counts = c->d_counts;
    a.ncl_max = c->ncl_max;
    a.dummy_slot = c->nslot_max;
    a.cell_slot = c->d_cell_slot;
    a.cell_count = c->d_cell_count;
    a.stencil = st;
    a.nstencil = nst;
    a.lcap = c->lcap;
    a.overflow = c->d_overflow;
    a.bbox_filter = !list_exact(c) && lmin_half > pcut + lskin + cdiag + 1e-3 ? 1 : 0;
    a.rebuild = rb;
    a.skin = lskin;
    a.max_delta = c->max_delta;
    a.xs_row = c->d_xs_row;
    a.xs_part = c->d_xs_part;
    a.xs_tcap = c->m * c->xs_maxdeg;
    a.m = c->m;
    a.jlanes = pair_lanes(c);
    a.molecule = c->d_mol;
    a.ex_row = c->d_ex_row;
    a.ex_partner = c->d_ex_partner;
    a.ex_cls = c->d_ex_cls;
    a.ntypes1 = c->ntypes + 1;
    a.lj_f = c->d_lj_f;
    a.lj_d = c->d_lj_d;
    a.atom_lj_f = c->d_atom_lj_f;
    a.atom_lj_d = c->d_atom_lj_d;
    a.force_fix = c->d_force;
    a.force_slot = c->d_force_slot;
    a.epart = c->d_epart + (long long)pi * c->ncl_max * PM_NE;
    a.want_virial = vir ? 1 : 0;
    a.want_energy = want_energy ? 1 : 0;
  }
  // the list buffers may have grown while the passes were set up
  for (auto& pl : pair_launch) {
    pl.first.jlist = c->d_jlist;
    pl.first.jmask = c->d_jmask;
    pl.first.jcount = c->d_jcount;
    pl.first.lcap = c->lcap;
    if (inner) {
      // the pair cutoff + the inner skin, padded as the list cutoff is
      const double cp = std::sqrt(pl.second.cut2) + prune_skin(c);
      const float cf = (float)(cp * (1.0 + 1e-5) + 1e-4);
      pl.first.inner = 1;
      pl.first.cut2_prune = cf * cf;
      pl.first.ilist = c->d_ilist;
      pl.first.imask = c->d_imask;
      pl.first.icount = c->d_icount;
    }
  }
  std::vector<std::pair<PmPmeLaunch, int>> pme_launch;   // (launch, handle)
  double const_coul = 0.0, const_vir = 0.0;
  for (auto& hr : pme_rec) {
    ForceDesc& f = c->forces[hr.first];
    long long off = hr.second;
    if (f.recip) {
      PmPmeLaunch a{};
      a.n = n;
      a.nx = f.nx; a.ny = f.ny; a.nz = f.nz; a.nzc = f.nz / 2 + 1;
      a.plane = pme_plane(c, f);
      a.order = f.order;
      for (int k = 0; k < 3; k++) {
        a.hr[k] = (k == 0 ? f.nx : k == 1 ? f.ny : f.nz) / c->box.L[k];
        a.Lr[k] = 1.0 / c->box.L[k];
      }
      a.q = c->d_q;
      a.slot_pos = c->d_slot_pos;
      a.atom_of_slot = c->d_atom_of_slot;
      a.counts = c->d_counts;
      a.nslot_max = c->nslot_max;
      a.grid_fix = c->d_grid_fix;
      a.grid_r = c->d_grid_r;
      a.grid_c = c->d_grid_c;
      a.bmod = c->d_bmod;
      a.expf_tab = c->d_expt;
      a.inv_pi_v = 1.0 / (kPi * (c->box.L[0] * c->box.L[1] * c->box.L[2]));
      a.fac = kPi * kPi / (f.alpha * f.alpha);
      a.tw = c->d_tw;
      a.err = c->d_overflow;
      a.prio = pme_priority(c);
      a.force_fix = c->d_force;
      a.epart = c->d_epart + off * PM_NE;
      a.want_virial = vir ? 1 : 0;
      a.want_energy = want_energy ? 1 : 0;
      a.theta = c->d_theta;
      // spread bricks: 2x2x2 cells per brick (fewer where the tile would not fit 48 KB of LDS)
      a.cell_slot = c->d_cell_slot;
      a.cell_count = c->d_cell_count;
      const int ng[3] = {f.nx, f.ny, f.nz};
      int br = 2;
      for (;;) {
        size_t pts = 1;
        for (int d = 0; d < 3; d++) {
          a.ncell[d] = c->box.n[d];
          a.brick[d] = std::min(br, c->box.n[d]);
          a.nbrick[d] = (c->box.n[d] + a.brick[d] - 1) / a.brick[d];
          a.cell_du[d] = c->realcell[d] * ng[d] / c->box.L[d];
          double u = (c->box.origin[d] - 0.5 * c->box.L[d]) / c->box.L[d];
          u -= std::floor(u);
          a.cell_lo_u[d] = u * ng[d];
          int t = (int)std::ceil(a.brick[d] * a.cell_du[d]) + f.order + 3;
          a.tile[d] = std::min(t, ng[d]);
          pts *= (size_t)a.tile[d];
        }
        if (pts * 8 <= 48 * 1024 || br == 1) break;
        br = 1;
      }
      if (a.brick[0] * a.brick[1] * a.brick[2] > 64)
        return fail(c, PM_ERR_UNSUPPORTED, "internal: spread brick of more than 64 cells");
      if (f.order < 2 || f.order > 12 || (f.order & 1))
        return fail(c, PM_ERR_UNSUPPORTED, "PME order %d not implemented on the GPU (even orders 2 to 12)", f.order);
      // spread fused with the z transform: an atom lies within half the
      // list skin of the position it was binned at (else this evaluation
      // rebuilds the layout), padded for the float displacement test
      const double drift_A = (reuse ? 0.5 * skin * (1.0 + 1e-3) : 0.0) + 1e-3;
      for (int d = 0; d < 3; d++) a.drift_u[d] = drift_A * ng[d] / c->box.L[d] + 0.01;
      a.spread_zfwd = !adj && c->pme_spread_option != 0 ? pm_pme_spread_zfwd_fits(a, c->precision) : 0;
      if (!a.spread_zfwd && (size_t)a.tile[0] * a.tile[1] * a.tile[2] * 8 > 64 * 1024)
        return fail(c, PM_ERR_UNSUPPORTED, "PME grid too fine for the cell size: spread tile exceeds 64 KB of LDS (use a smaller cellsize)");
      pme_launch.push_back(std::make_pair(a, hr.first));
    }
    if (f.corr) {
      // point self energy and charged-system term (:604-628)
      double q2 = 0.0, qs = 0.0;
      for (int i = 0; i < n; i++) q2 += c->q[i] * c->q[i];
      for (int i = 0; i < n; i++) qs += c->q[i];
      const_coul += -q2 * f.alpha / std::sqrt(kPi);
      if (std::fabs(qs * 0.00268283) > 1.0e-5) {
        double V = c->box.L[0] * c->box.L[1] * c->box.L[2];
        double ch = -kPi / (2.0 * V * f.alpha * f.alpha) * qs * qs;
        const_coul += ch;
        if (vir) const_vir += ch;
      }
    }
  }
  if (want_energy && !c->d_ered &&
      (r = dalloc(c, &c->d_ered, (size_t)pm_reduce_scratch_records((int)rec) * PM_NE)))
    return r;

  if (adj) {
    // adjoint kernels, all on the first stream
    pm_launch_slot_vec(s, c->d_counts, c->d_atom_of_slot, adj->mode == 1 ? adj->v : c->d_scratch3, adj->slot_v,
                       c->nslot_max);
    pm_launch_zero_ll(s, adj->slot, 3LL * (c->nslot_max + 1));
    pm_launch_zero_ll(s, adj->fix, 3LL * n);
    if (adj->mode == 1) pm_launch_zero_d(s, adj->hv, 3LL * n);
    else pm_launch_zero_d(s, adj->dq, n);
    for (auto& pl : pair_launch) {
      pm_launch_list(s, c->box, pl.first, pl.second);
      pm_launch_adjoint(s, c->box, pl.first, pl.second, c->precision, adj->mode, adj->slot_v, adj->slot);
    }
    if (!pair_launch.empty())
      pm_launch_fold_slots(s, c->nslot_max, c->d_counts, c->d_atom_of_slot, adj->slot, adj->fix, nullptr, nullptr, 0);
    for (auto& pl : pme_launch)
      pm_launch_pme_adjoint(s, pl.first, c->plans, c->d_slot_pos64, c->precision, adj->mode == 1 ? 1 : 0,
                            adj->slot_v, adj->mode == 1 ? adj->hv : adj->dq);
    for (auto& hr : pme_rec) {
      ForceDesc& f = c->forces[hr.first];
      if (!f.corr) continue;
      pm_launch_corr_adjoint(s, c->box, n, c->d_pos, c->d_q, c->d_cx_row, c->d_cx_partner, c->d_cx_cls, f.alpha,
                             c->csf, adj->mode == 1 ? 1 : 0, adj->v, adj->mode == 1 ? adj->hv : adj->dq);
    }
    HIPCHECK(c, hipGetLastError());
    return PM_OK;
  }

  // fork: PME, correction and bonded terms on the second stream, the pair
  // lists and pair kernels on the first; they write different force arrays
  // (per atom, per slot) and different energy records
  hipStream_t s2 = c->stream2;
  // the second stream pays off beside the PME kernels; bonded terms alone
  // stay on the first
  const bool side = !pme_rec.empty();
  if (!side) s2 = s;
  if (side) {
    HIPCHECK(c, hipEventRecord(c->ev_fork, s));
    HIPCHECK(c, hipStreamWaitEvent(s2, c->ev_fork, 0));
  }
  PmBondedLaunch b{};
  if (any_bonded) {
    b.n = n;
    b.pos = c->d_pos;
    b.b_row = c->d_b_row; b.b_idx = c->d_b_idx; b.a_row = c->d_a_row; b.a_idx = c->d_a_idx;
    b.nb = c->nbonds; b.na = c->nangles;
    b.b_a = c->d_b_a; b.b_b = c->d_b_b; b.b_k = c->d_b_k; b.b_r0 = c->d_b_r0;
    b.a_a = c->d_a_a; b.a_b = c->d_a_b; b.a_c = c->d_a_c; b.a_k = c->d_a_k; b.a_t0 = c->d_a_t0;
    b.a_kub = c->d_a_kub; b.a_rub = c->d_a_rub;
    b.do_bond = do_bond ? 1 : 0;
    b.do_angle = do_angle ? 1 : 0;
    for (int k = 0; k < 2; k++) {
      b.t_row[k] = c->d_t_row[k]; b.t_idx[k] = c->d_t_idx[k]; b.nt[k] = c->ntors[k];
      b.t_atoms[k] = c->d_t_atoms[k]; b.t_term[k] = c->d_t_term[k];
      b.t_k[k] = c->d_t_k[k]; b.t_n[k] = c->d_t_n[k]; b.t_phase[k] = c->d_t_phase[k];
      b.do_tors[k] = do_tors[k] ? 1 : 0;
    }
    b.force_fix = c->d_force;
    b.epart = c->d_epart + bonded_off * PM_NE;
    b.want_virial = vir ? 1 : 0;
    b.want_energy = want_energy ? 1 : 0;
  }
  // with one PME reciprocal term, the gather kernel's thread for an atom
  // also adds that atom's correction and bonded forces (two launches fewer)
  PmAtomTail tail{};
  const bool fuse = pme_launch.size() == 1;
  if (fuse) {
    const ForceDesc& f = c->forces[pme_launch[0].second];
    tail.box = c->box;
    if (f.corr) {
      tail.corr = 1;
      tail.pos = c->d_pos; tail.q = c->d_q;
      tail.row = c->d_cx_row; tail.partner = c->d_cx_partner; tail.cls = c->d_cx_cls;
      tail.alpha = f.alpha; tail.csf = c->csf;
    }
    if (any_bonded) {
      tail.bonded = 1;
      tail.b = b;
    }
  }
  for (auto& pl : pme_launch) {
    TimedScope ts(c, PM_T_PME, s2);
    pm_launch_pme_plans(s2, pl.first, c->plans, c->d_slot_pos64, c->precision, fuse ? &tail : nullptr);
  }
  for (auto& hr : pme_rec) {
    ForceDesc& f = c->forces[hr.first];
    if (!f.corr) continue;
    const bool in_tail = fuse && hr.first == pme_launch[0].second;
    long long off = hr.second + (f.recip ? pm_pme_xpass_blocks(f.nx, f.ny, f.nz, c->precision, pme_plane(c, f)) : 0);
    int ntab = (int)c->tab_a.size();
    int nblk = (ntab + 255) / 256;
    if (in_tail && !want_energy) continue;
    TimedScope ts(c, PM_T_CORR, s2);
    pm_launch_correction(s2, c->box, in_tail ? 0 : n, c->d_pos, c->d_q, c->d_cx_row, c->d_cx_partner, c->d_cx_cls,
                         f.alpha, c->csf, vir ? 1 : 0, c->d_force, ntab, c->d_tab_a, c->d_tab_b, c->d_tab_c,
                         c->d_epart + off * PM_NE, want_energy ? nblk : 0);
  }
  for (auto& hr : pme_rec) {
    ForceDesc& f = c->forces[hr.first];
    if (!f.real || !pme_lattice(c, f)) continue;
    long long off = hr.second + (f.recip ? pm_pme_xpass_blocks(f.nx, f.ny, f.nz, c->precision, pme_plane(c, f)) : 0) +
                    (f.corr ? ((long long)c->tab_a.size() + 255) / 256 : 0);
    TimedScope ts(c, PM_T_PME, s2);
    pm_launch_ewald_lattice(s2, c->box, n, c->d_pos, c->d_q, c->d_xs_row, c->d_xs_part, c->d_lat, c->nlat,
                            f.rc * f.rc, f.alpha, c->csf, vir ? 1 : 0, c->d_force, c->d_epart + off * PM_NE,
                            lattice_blocks(n));
  }
  if (any_bonded && !(fuse && !want_energy)) {
    PmBondedLaunch bl = b;
    if (fuse) bl.n = 0;   // the forces came with the gather: energies only
    TimedScope ts(c, PM_T_BONDED, s2);
    pm_launch_bonded(s2, c->box, bl);
  }
  if (side) HIPCHECK(c, hipEventRecord(c->ev_join, s2));

  for (auto& pl : pair_launch) {
    {
      TimedScope ts(c, PM_T_LIST);
      pm_launch_list(s, c->box, pl.first, pl.second);
      if (pl.first.inner) pm_launch_prune(s, c->box, pl.first);
    }
    {
      TimedScope ts(c, PM_T_PAIR);
      pm_launch_pair(s, c->box, pl.first, pl.second, c->precision);
    }
  }
  // join, then the per-slot pair forces into the per-atom forces
  if (side) HIPCHECK(c, hipStreamWaitEvent(s, c->ev_join, 0));
  if (!passes.empty()) {
    TimedScope ts(c, PM_T_BIN);
    // with reuse also the end of the evaluation: the flag's readers all ran
    pm_launch_fold_slots(s, c->nslot_max, c->d_counts, c->d_atom_of_slot, c->d_force_slot, c->d_force,
                         reuse ? c->d_rebuild : nullptr, c->d_stats + 2, 1);
  }
  if (want_energy) {
    TimedScope ts(c, PM_T_REDUCE);
    pm_launch_reduce_energy(s, c->d_epart, (int)rec, const_coul, const_vir, c->d_energy, 1, c->d_ered);
  }
  (void)any_pme_recip;
  HIPCHECK(c, hipGetLastError());
  return PM_OK;
}

int check_overflow(pm_context* c) {
  if (!c->d_overflow) return PM_OK;
  int ov = 0;
  HIPCHECK(c, hipMemcpyAsync(&ov, c->d_overflow, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (ov) {
    HIPCHECK(c, hipMemset(c->d_overflow, 0, sizeof(int)));
    if (ov & PM_FLAG_NONFINITE)
      return fail(c, PM_ERR_NUMERIC, "a position is not finite (or lies more than 2^31 box lengths away): the "
                                     "forces of this evaluation are not valid");
    if (ov & PM_FLAG_FORCE_RANGE)
      return fail(c, PM_ERR_NUMERIC, "a force term is not finite or exceeds the fixed-point range of "
                                     "+-2^23 kcal/mol/A (overlapping atoms?): the forces of this evaluation are not valid");
    return fail(c, PM_ERR_CAPACITY, "pair list capacity (%d entries per cluster) exceeded: atoms are packed far denser than the box average", c->lcap);
  }
  return PM_OK;
}

// list statistics of the last evaluation (on demand: one reduction + copy)
int list_stats(pm_context* c) {
  // entries, lanes (of the inner lists when the last pass kept them), list
  // builds and inner-list prunes without a build (both cumulative)
  unsigned long long st[4] = {0, 0, 0, 0};
  int counts[4] = {0, 0, 0, 0};
  if (c->d_stats && c->d_jcount && c->d_counts) {
    const int* len = c->inner_last && c->d_icount ? c->d_icount : c->d_jcount;
    pm_launch_list_stats(c->stream, len, c->d_counts, c->d_stats, c->m, pair_lanes(c));
    HIPCHECK(c, hipMemcpyAsync(st, c->d_stats, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(counts, c->d_counts, sizeof(counts), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  }
  c->last_stats[0] = (long long)st[0];
  c->last_stats[1] = (long long)st[1];
  c->last_stats[2] = counts[1];
  c->last_stats[3] = counts[0];
  c->last_stats[4] = (long long)st[2];
  c->last_stats[5] = (long long)st[3];
  return PM_OK;
}

// The SHAKE/RATTLE constraint list of buildRattleShakeBondConstraintList
// (topology/topologyutilities.cpp:782-865): every bond at its rest length,
// then for every angle whose end atoms both weigh less than 5 the end-to-end
// distance sqrt(b^2 + c^2 - 2 b c cos(theta0)) of the first pair of bonds
// (a1's bonds in bond order, then a3's) joining the end atoms to the vertex.
// The constraints are then ordered by connected group of atoms (groups in
// order of their first constraint), each group's in list order, and the
// group arrays uploaded.
int build_constraints(pm_context* c) {
  if (c->cons_built) return PM_OK;
  const int n = c->n, nb = (int)c->h_ba.size(), na = (int)c->h_aa.size();
  std::vector<int> ca1, ca2;
  std::vector<double> cr0;
  for (int k = 0; k < nb; k++) { ca1.push_back(c->h_ba[k]); ca2.push_back(c->h_bb[k]); cr0.push_back(c->h_br0[k]); }
  std::vector<std::vector<int>> mybonds(n);
  for (int k = 0; k < nb; k++) { mybonds[c->h_ba[k]].push_back(k); mybonds[c->h_bb[k]].push_back(k); }
  auto sorted = [](int a, int b) { return a < b ? std::make_pair(a, b) : std::make_pair(b, a); };
  for (int i = 0; i < na; i++) {
    const int a1 = c->h_aa[i], a2 = c->h_ab[i], a3 = c->h_ac[i];
    if (!(c->mass[a1] < 5 && c->mass[a3] < 5)) continue;
    const auto p1 = sorted(a1, a2), p2 = sorted(a2, a3);
    int b1 = -1, b2 = -1;
    for (int j : mybonds[a1]) {
      if (sorted(c->h_ba[j], c->h_bb[j]) != p1) continue;
      for (int k : mybonds[a3]) {
        if (sorted(c->h_ba[k], c->h_bb[k]) != p2) continue;
        if (b1 < 0 && b2 < 0) { b1 = j; b2 = k; }
      }
    }
    if (b1 > -1 && b2 > -1) {
      const double b = c->h_br0[b1], cc = c->h_br0[b2], alpha = c->h_at0[i];
      ca1.push_back(a1); ca2.push_back(a3); cr0.push_back(std::sqrt(b * b + cc * cc - 2 * b * cc * std::cos(alpha)));
    }
  }
  const int m = (int)ca1.size();
  // connected groups (union-find over the constraint graph)
  std::vector<int> par(n);
  for (int i = 0; i < n; i++) par[i] = i;
  auto root = [&](int x) { while (par[x] != x) { par[x] = par[par[x]]; x = par[x]; } return x; };
  for (int k = 0; k < m; k++) { const int ra = root(ca1[k]), rb = root(ca2[k]); if (ra != rb) par[ra] = rb; }
  std::vector<int> gid(n, -1), grp(m);
  int ng = 0;
  for (int k = 0; k < m; k++) {
    const int r = root(ca1[k]);
    if (gid[r] < 0) gid[r] = ng++;
    grp[k] = gid[r];
  }
  std::vector<int> gstart(ng + 1, 0), a1s(m), a2s(m);
  std::vector<double> r0s(m);
  for (int k = 0; k < m; k++) gstart[grp[k] + 1]++;
  for (int g = 0; g < ng; g++) gstart[g + 1] += gstart[g];
  {
    std::vector<int> fill(gstart.begin(), gstart.end() - 1);
    for (int k = 0; k < m; k++) { const int at = fill[grp[k]]++; a1s[at] = ca1[k]; a2s[at] = ca2[k]; r0s[at] = cr0[k]; }
  }
  // groups of at most 4 atoms and 6 constraints: their atoms (in order of
  // appearance) and each constraint's atoms as local indices
  std::vector<int4> gatom(std::max(ng, 1), make_int4(-1, -1, -1, -1));
  std::vector<int> la(std::max(m, 1), 0);
  for (int g = 0; g < ng; g++) {
    int at[4] = {-1, -1, -1, -1}, na_ = 0;
    bool small = true;
    auto local = [&](int a) {
      for (int i = 0; i < na_; i++) if (at[i] == a) return i;
      if (na_ == 4) { small = false; return 0; }
      at[na_] = a;
      return na_++;
    };
    for (int k = gstart[g]; k < gstart[g + 1] && small; k++) {
      const int l1 = local(a1s[k]), l2 = local(a2s[k]);
      la[k] = l1 | (l2 << 2);
    }
    if (small && gstart[g + 1] - gstart[g] <= 6) {
      gatom[g] = make_int4(at[0], at[1], at[2], at[3]);
      const int k = gstart[g];
      if (na_ == 3 && gstart[g + 1] - k == 3 && la[k] == (0 | 1 << 2) && la[k + 1] == (0 | 2 << 2) &&
          la[k + 2] == (1 | 2 << 2))
        gatom[g].w = PM_CONS_WATER;
    }
  }
  std::vector<double> rm(n);
  for (int i = 0; i < n; i++) rm[i] = 1 / c->mass[i];
  int r;
  if ((r = upload(c, &c->d_cons_gstart, gstart))) return r;
  if ((r = upload(c, &c->d_cons_la, la))) return r;
  if ((r = upload(c, &c->d_cons_gatom, gatom))) return r;
  if ((r = upload(c, &c->d_cons_a1, a1s))) return r;
  if ((r = upload(c, &c->d_cons_a2, a2s))) return r;
  if ((r = upload(c, &c->d_cons_r0, r0s))) return r;
  if ((r = upload(c, &c->d_rmass, rm))) return r;
  if (!c->d_cons_sweeps && (r = dalloc(c, &c->d_cons_sweeps, 1))) return r;
  if ((r = dalloc(c, &c->d_cons_scratch, 3 * (size_t)n))) return r;
  if ((r = dalloc(c, &c->d_shake_last, 3 * (size_t)n))) return r;
  c->cons_n = m;
  c->cons_ngroup = ng;
  c->cons_err_cap = 0;
  c->cons_built = true;
  return PM_OK;
}

// per-sweep error buffers for up to max_iter + 1 sweep counts
int ensure_constraint_err(pm_context* c, int max_iter) {
  if (max_iter + 1 <= c->cons_err_cap && c->d_cons_err) return PM_OK;
  int r;
  const int cap = max_iter + 1;
  // wave sums: four waves per block of 256 groups
  if ((r = dalloc(c, &c->d_cons_err, (size_t)cap * 4 * std::max((c->cons_ngroup + 255) / 256, 1)))) return r;
  c->cons_err_cap = cap;
  return PM_OK;
}

PmConstraints constraint_launch(pm_context* c, double eps, int max_iter, double dt_fs) {
  PmConstraints k{};
  k.ngroup = c->cons_ngroup;
  k.ncons = c->cons_n;
  k.maxit = max_iter;
  k.gstart = c->d_cons_gstart;
  k.a1 = c->d_cons_a1;
  k.a2 = c->d_cons_a2;
  k.la = c->d_cons_la;
  k.gatom = c->d_cons_gatom;
  k.r0 = c->d_cons_r0;
  k.rmass = c->d_rmass;
  k.eps = eps;
  k.dt = dt_fs / kTimeFactor;   // getTimestep() / Constant::TIMEFACTOR (ModifierShake.cpp:27)
  k.err = c->d_cons_err;
  k.sweeps = c->d_cons_sweeps;
  k.scratch = c->d_cons_scratch;
  return k;
}

void free_all(pm_context* c) {
  dfree(c->d_q); dfree(c->d_mass); dfree(c->d_type); dfree(c->d_mol);
  dfree(c->d_ex_row); dfree(c->d_ex_partner); dfree(c->d_ex_cls);
  dfree(c->d_xs_row); dfree(c->d_xs_part); dfree(c->d_jmask);
  dfree(c->d_atom_lj_f); dfree(c->d_atom_lj_d);
  c->lj_R.clear(); c->lj_e.clear();
  dfree(c->d_adj_sv); dfree(c->d_adj_slot); dfree(c->d_adj_fix); dfree(c->d_adj_v); dfree(c->d_adj_out);
  c->adj_nslot = 0; c->adj_n = 0;
  dfree(c->d_cx_row); dfree(c->d_cx_partner); dfree(c->d_cx_cls);
  dfree(c->d_tab_a); dfree(c->d_tab_b); dfree(c->d_tab_c);
  dfree(c->d_lj_f); dfree(c->d_lj_d);
  dfree(c->d_pos); dfree(c->d_vel); dfree(c->d_force); dfree(c->d_energy); dfree(c->d_scratch3);
  dfree(c->d_cell_of); dfree(c->d_key); dfree(c->d_cell_count); dfree(c->d_bin_count); dfree(c->d_cell_start); dfree(c->d_cell_slot);
  dfree(c->d_scan_aux);
  dfree(c->d_xbuild); dfree(c->d_rebuild); c->list_stale = true;
  dfree(c->d_xprune); dfree(c->d_ilist); dfree(c->d_imask); dfree(c->d_icount); c->ilist_alloc = 0;
  dfree(c->d_cell_rank); dfree(c->d_force_slot); dfree(c->d_cell_atoms); dfree(c->d_cluster_cell); dfree(c->d_counts);
  dfree(c->d_slot_pos); dfree(c->d_slot_pos64); dfree(c->d_slot_meta); dfree(c->d_atom_of_slot);
  dfree(c->d_jlist); dfree(c->d_jcount); dfree(c->d_stats); dfree(c->d_epart); dfree(c->d_ered);
  dfree(c->d_grid_fix);
  if (c->d_grid_r) hipFree(c->d_grid_r);
  if (c->d_grid_c) hipFree(c->d_grid_c);
  if (c->d_theta) hipFree(c->d_theta);
  c->d_grid_r = c->d_grid_c = c->d_theta = nullptr;
  dfree(c->d_lat);
  c->nlat = 0;
  dfree(c->d_bmod); dfree(c->d_expt); dfree(c->d_tw);
  dfree(c->d_b_a); dfree(c->d_b_b); dfree(c->d_b_row); dfree(c->d_b_idx); dfree(c->d_b_k); dfree(c->d_b_r0);
  dfree(c->d_a_a); dfree(c->d_a_b); dfree(c->d_a_c); dfree(c->d_a_row); dfree(c->d_a_idx);
  dfree(c->d_a_k); dfree(c->d_a_t0); dfree(c->d_a_kub); dfree(c->d_a_rub);
  c->nbonds = c->nangles = 0;
  for (int k = 0; k < 2; k++) {
    dfree(c->d_t_row[k]); dfree(c->d_t_idx[k]); dfree(c->d_t_term[k]); dfree(c->d_t_atoms[k]);
    dfree(c->d_t_k[k]); dfree(c->d_t_phase[k]); dfree(c->d_t_n[k]);
    c->ntors[k] = 0;
  }
  dfree(c->d_cons_gstart); dfree(c->d_cons_a1); dfree(c->d_cons_a2); dfree(c->d_cons_sweeps);
  dfree(c->d_cons_r0); dfree(c->d_rmass); dfree(c->d_cons_err); dfree(c->d_cons_la); dfree(c->d_cons_gatom);
  dfree(c->d_cons_scratch); dfree(c->d_shake_last);
  dfree(c->d_cons_x); dfree(c->d_cons_v); dfree(c->d_cons_xl);
  c->cons_built = false;
  c->cons_n = c->cons_ngroup = c->cons_err_cap = 0;
  free_stencils(c);
  c->cells_alloc = 0;
  c->jlist_alloc = 0;
  c->epart_alloc = 0;
  c->pme_nx = c->pme_ny = c->pme_nz = 0;
  c->pme_prec = -1;
}

int set_device(pm_context* c) {
  HIPCHECK(c, hipSetDevice(c->device));
  return PM_OK;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int pm_abi_version(void) { return PM_ABI_VERSION; }

const char* pm_last_error(const pm_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pm_create(int device, pm_context** out) {
  if (!out) return PM_ERR_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return PM_ERR_DEVICE;
  if (device < 0 || device >= count) return PM_ERR_ARGUMENT;
  pm_context* c = new pm_context();
  c->device = device;
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
      c->lds_max = v;
  }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return PM_ERR_DEVICE;
  }
  // both streams at the default priority
  if (hipStreamCreate(&c->stream2) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    if (c->stream2) hipStreamDestroy(c->stream2);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    hipStreamDestroy(c->stream);
    delete c;
    return PM_ERR_DEVICE;
  }
  // the error flag word (PM_FLAG_*) of every kernel of this context: lives
  // as long as the context, so the box's pointer to it never dangles
  if (hipMalloc((void**)&c->d_overflow, sizeof(int)) != hipSuccess ||
      hipMemset(c->d_overflow, 0, sizeof(int)) != hipSuccess) {
    if (c->d_overflow) hipFree(c->d_overflow);
    hipEventDestroy(c->ev_fork);
    hipEventDestroy(c->ev_join);
    hipStreamDestroy(c->stream2);
    hipStreamDestroy(c->stream);
    delete c;
    return PM_ERR_DEVICE;
  }
  c->box.err = c->d_overflow;
  // the allocation-free entry point's work space on this device (its address
  // is looked up once per device, here, with the device current)
  if (pm_scratch_init(device) != 0) {
    pm_destroy(c);
    return PM_ERR_DEVICE;
  }
  *out = c;
  return PM_OK;
}

void pm_destroy(pm_context* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  hipStreamSynchronize(ctx->stream2);
  harvest_timing(ctx);
  for (hipEvent_t e : ctx->ev_pool) hipEventDestroy(e);
  drop_graph(ctx);
  free_all(ctx);
  dfree(ctx->d_md);
  dfree(ctx->d_overflow);
  hipEventDestroy(ctx->ev_fork);
  hipEventDestroy(ctx->ev_join);
  hipStreamDestroy(ctx->stream2);
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

int pm_set_precision(pm_context* c, int precision) {
  if (c) c->list_stale = true;
  if (!c) return PM_ERR_ARGUMENT;
  if (precision != PM_PRECISION_MIXED && precision != PM_PRECISION_DOUBLE)
    return fail(c, PM_ERR_ARGUMENT, "precision must be PM_PRECISION_MIXED (0) or PM_PRECISION_DOUBLE (1)");
  c->precision = precision;
  c->md_forces_valid = false;
  c->gen++;
  return PM_OK;
}

int pm_set_topology(pm_context* c, int n, const double* q, const int* type, int ntypes, const double* ljA,
                    const double* ljB, const double* ljA14, const double* ljB14, const int* molecule,
                    const double* mass, int nexcl, const int* ex_a, const int* ex_b, const int* ex_class,
                    double csf) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (n <= 0 || !q || !type || ntypes <= 0 || !ljA || !ljB || !ljA14 || !ljB14 || !molecule)
    return fail(c, PM_ERR_ARGUMENT, "pm_set_topology: n > 0 and non-null arrays required");
  if (n >= (1 << 23) - 1) return fail(c, PM_ERR_CAPACITY, "pm_set_topology: at most %d atoms", (1 << 23) - 2);
  // the pair kernels stage the (types + 1)^2 parameter table in LDS: at most
  // 44 types keep the FP64 table (45^2 x 32 bytes) within 64 KiB
  if (ntypes > 44) return fail(c, PM_ERR_UNSUPPORTED, "pm_set_topology: at most 44 atom types (got %d)", ntypes);
  if (nexcl > 0 && (!ex_a || !ex_b || !ex_class)) return fail(c, PM_ERR_ARGUMENT, "pm_set_topology: exclusion arrays missing");
  for (int i = 0; i < n; i++)
    if (type[i] < 0 || type[i] >= ntypes) return fail(c, PM_ERR_ARGUMENT, "atom %d has type %d outside [0, %d)", i, type[i], ntypes);
  for (int k = 0; k < nexcl; k++)
    if (ex_a[k] < 0 || ex_a[k] >= n || ex_b[k] < 0 || ex_b[k] >= n || ex_class[k] < 0 || ex_class[k] > 2)
      return fail(c, PM_ERR_ARGUMENT, "exclusion %d (%d, %d, class %d) out of range", k, ex_a[k], ex_b[k], ex_class[k]);
  free_all(c);
  c->n = n;
  c->ntypes = ntypes;
  c->q.assign(q, q + n);
  c->type.assign(type, type + n);
  c->molecule.assign(molecule, molecule + n);
  c->mass.assign(n, 1.0);
  if (mass) c->mass.assign(mass, mass + n);
  c->ljA.assign(ljA, ljA + ntypes * ntypes);
  c->ljB.assign(ljB, ljB + ntypes * ntypes);
  c->ljA14.assign(ljA14, ljA14 + ntypes * ntypes);
  c->ljB14.assign(ljB14, ljB14 + ntypes * ntypes);
  c->csf = csf;
  c->ex_a.assign(ex_a, ex_a + nexcl);
  c->ex_b.assign(ex_b, ex_b + nexcl);
  c->ex_c.assign(ex_class, ex_class + nexcl);
  // lookup CSR by lower atom in add order (ExclusionTable::add/check, last add wins)
  std::vector<int> row(n + 1, 0), partner(nexcl), cls(nexcl);
  c->max_delta = 0;
  for (int k = 0; k < nexcl; k++) row[std::min(ex_a[k], ex_b[k]) + 1]++;
  for (int i = 0; i < n; i++) row[i + 1] += row[i];
  {
    std::vector<int> fill(row.begin(), row.end() - 1);
    for (int k = 0; k < nexcl; k++) {
      int a = std::min(ex_a[k], ex_b[k]), b = std::max(ex_a[k], ex_b[k]);
      partner[fill[a]] = b;
      cls[fill[a]] = ex_class[k];
      fill[a]++;
      c->max_delta = std::max(c->max_delta, b - a);
    }
  }
  // resolved exclusion of every pair as OneAtomPair sees it: the class of the
  // last add (ExclusionTable::add stores it, NONE included), consulted only
  // for atoms of one molecule (OneAtomPair.h:149-153).  Stored for both atoms,
  // partner | class << 24, sorted by partner.
  {
    std::map<std::pair<int, int>, int> last;
    for (int k = 0; k < nexcl; k++) {
      int a = std::min(ex_a[k], ex_b[k]), b = std::max(ex_a[k], ex_b[k]);
      if (a != b) last[std::make_pair(a, b)] = ex_class[k];
    }
    std::vector<std::vector<int>> adj(n);
    for (auto& kv : last) {
      int a = kv.first.first, b = kv.first.second;
      if (kv.second == PM_EXCL_NONE || molecule[a] != molecule[b]) continue;
      adj[a].push_back(b | (kv.second << 24));
      adj[b].push_back(a | (kv.second << 24));
    }
    std::vector<int> xrow(n + 1, 0), xpart;
    c->xs_maxdeg = 0;
    for (int i = 0; i < n; i++) {
      std::sort(adj[i].begin(), adj[i].end(), [](int u, int v) { return (u & 0xffffff) < (v & 0xffffff); });
      xrow[i + 1] = xrow[i] + (int)adj[i].size();
      c->xs_maxdeg = std::max(c->xs_maxdeg, (int)adj[i].size());
      xpart.insert(xpart.end(), adj[i].begin(), adj[i].end());
    }
    if (c->xs_maxdeg > 256) {
      free_all(c);
      return fail(c, PM_ERR_UNSUPPORTED, "pm_set_topology: an atom has %d exclusion partners (at most 256)", c->xs_maxdeg);
    }
    if ((r = upload(c, &c->d_xs_row, xrow))) return r;
    if ((r = upload(c, &c->d_xs_part, xpart))) return r;
  }
  // the table (ExclusionTable::myTable): the first add with class != NONE and
  // a != b of every pair, in add order (mySet orders pairs by (a, b) only,
  // topology/ExclusionTable.cpp:20-28,112-121)
  c->tab_a.clear(); c->tab_b.clear(); c->tab_c.clear();
  {
    std::set<std::pair<int, int>> seen;
    for (int k = 0; k < nexcl; k++) {
      int a = std::min(ex_a[k], ex_b[k]), b = std::max(ex_a[k], ex_b[k]);
      if (ex_class[k] == PM_EXCL_NONE || a == b) continue;
      auto t = std::make_pair(a, b);
      if (seen.count(t)) continue;
      seen.insert(t);
      c->tab_a.push_back(a); c->tab_b.push_back(b); c->tab_c.push_back(ex_class[k]);
    }
  }
  // correction CSR per atom: partner*2 + (1 if this atom is the pair's second atom)
  int ntab = (int)c->tab_a.size();
  std::vector<int> crow(n + 1, 0), cpart(2 * ntab), ccls(2 * ntab);
  for (int t = 0; t < ntab; t++) { crow[c->tab_a[t] + 1]++; crow[c->tab_b[t] + 1]++; }
  for (int i = 0; i < n; i++) crow[i + 1] += crow[i];
  {
    std::vector<int> fill(crow.begin(), crow.end() - 1);
    for (int t = 0; t < ntab; t++) {
      int a = c->tab_a[t], b = c->tab_b[t];
      cpart[fill[a]] = 2 * b; ccls[fill[a]] = c->tab_c[t]; fill[a]++;
      cpart[fill[b]] = 2 * a + 1; ccls[fill[b]] = c->tab_c[t]; fill[b]++;
    }
  }
  // LJ tables with one extra all-zero type for padding slots
  int t1 = ntypes + 1;
  std::vector<float4> ljf(t1 * t1, make_float4(0, 0, 0, 0));
  std::vector<double4> ljd(t1 * t1, make_double4(0, 0, 0, 0));
  for (int a = 0; a < ntypes; a++)
    for (int b = 0; b < ntypes; b++) {
      int k = a * ntypes + b;
      ljf[a * t1 + b] = make_float4((float)ljA[k], (float)ljB[k], (float)ljA14[k], (float)ljB14[k]);
      ljd[a * t1 + b] = make_double4(ljA[k], ljB[k], ljA14[k], ljB14[k]);
    }
  std::vector<double> mass_v = c->mass;
  if ((r = upload(c, &c->d_q, c->q))) return r;
  if ((r = upload(c, &c->d_mass, mass_v))) return r;
  if ((r = upload(c, &c->d_type, c->type))) return r;
  if ((r = upload(c, &c->d_mol, c->molecule))) return r;
  if ((r = upload(c, &c->d_ex_row, row))) return r;
  if ((r = upload(c, &c->d_ex_partner, partner))) return r;
  if ((r = upload(c, &c->d_ex_cls, cls))) return r;
  if ((r = upload(c, &c->d_cx_row, crow))) return r;
  if ((r = upload(c, &c->d_cx_partner, cpart))) return r;
  if ((r = upload(c, &c->d_cx_cls, ccls))) return r;
  if ((r = upload(c, &c->d_tab_a, c->tab_a))) return r;
  if ((r = upload(c, &c->d_tab_b, c->tab_b))) return r;
  if ((r = upload(c, &c->d_tab_c, c->tab_c))) return r;
  if ((r = upload(c, &c->d_lj_f, ljf))) return r;
  if ((r = upload(c, &c->d_lj_d, ljd))) return r;
  if ((r = dalloc(c, &c->d_pos, 3 * (size_t)n))) return r;
  if ((r = dalloc(c, &c->d_vel, 3 * (size_t)n))) return r;
  HIPCHECK(c, hipMemset(c->d_vel, 0, 3 * (size_t)n * sizeof(double)));
  if ((r = dalloc(c, &c->d_force, 3 * (size_t)n))) return r;
  if ((r = dalloc(c, &c->d_scratch3, 3 * (size_t)n))) return r;
  if ((r = dalloc(c, &c->d_energy, PM_NE))) return r;
  HIPCHECK(c, hipMemset(c->d_energy, 0, PM_NE * sizeof(double)));
  for (auto& f : c->forces) f.cached = false;
  c->have_topo = true;
  c->md_forces_valid = false;
  c->gen++;
  return PM_OK;
}

int pm_set_lj_per_atom(pm_context* c, const double* sigma, const double* epsilon, const double* sigma14,
                       const double* epsilon14) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (!c->have_topo) return fail(c, PM_ERR_STATE, "no topology: call pm_set_topology first");
  c->gen++;
  c->md_forces_valid = false;
  if (!sigma) {
    dfree(c->d_atom_lj_f);
    dfree(c->d_atom_lj_d);
    c->lj_R.clear();
    c->lj_e.clear();
    return PM_OK;
  }
  if (!epsilon) return fail(c, PM_ERR_ARGUMENT, "pm_set_lj_per_atom: epsilon missing");
  const int n = c->n;
  const double rfac = std::pow(2.0, -5.0 / 6.0);   // Rmin/2 = sigma 2^(1/6) / 2
  std::vector<float4> f(n);
  std::vector<double4> d(n);
  c->lj_R.assign(n, 0.0);
  c->lj_e.assign(n, 0.0);
  for (int i = 0; i < n; i++) {
    const double s1 = sigma[i], e1 = epsilon[i];
    const double s4 = sigma14 ? sigma14[i] : s1, e4 = epsilon14 ? epsilon14[i] : e1;
    if (!(s1 >= 0) || !(e1 >= 0) || !(s4 >= 0) || !(e4 >= 0))
      return fail(c, PM_ERR_ARGUMENT, "atom %d: sigma and epsilon must be >= 0", i);
    const double R1 = s1 * rfac, E1 = std::sqrt(e1), R4 = s4 * rfac, E4 = std::sqrt(e4);
    c->lj_R[i] = R1;
    c->lj_e[i] = E1;
    d[i] = make_double4(R1, E1, R4, E4);
    f[i] = make_float4((float)R1, (float)E1, (float)R4, (float)E4);
  }
  if ((r = upload(c, &c->d_atom_lj_f, f))) return r;
  if ((r = upload(c, &c->d_atom_lj_d, d))) return r;
  return PM_OK;
}

int pm_set_bonded(pm_context* c, int nb, const int* ba, const int* bb, const double* bk, const double* br0,
                  int na, const int* aa, const int* ab, const int* ac, const double* ak, const double* at0,
                  const double* akub, const double* arub) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (!c->have_topo) return fail(c, PM_ERR_STATE, "no topology: call pm_set_topology first");
  if (nb < 0 || na < 0 || (nb > 0 && (!ba || !bb || !bk || !br0)) ||
      (na > 0 && (!aa || !ab || !ac || !ak || !at0 || !akub || !arub)))
    return fail(c, PM_ERR_ARGUMENT, "pm_set_bonded: arrays missing");
  const int n = c->n;
  for (int k = 0; k < nb; k++)
    if (ba[k] < 0 || ba[k] >= n || bb[k] < 0 || bb[k] >= n || ba[k] == bb[k])
      return fail(c, PM_ERR_ARGUMENT, "bond %d (%d, %d) out of range", k, ba[k], bb[k]);
  for (int k = 0; k < na; k++)
    if (aa[k] < 0 || aa[k] >= n || ab[k] < 0 || ab[k] >= n || ac[k] < 0 || ac[k] >= n ||
        aa[k] == ab[k] || ab[k] == ac[k] || aa[k] == ac[k])
      return fail(c, PM_ERR_ARGUMENT, "angle %d (%d, %d, %d) out of range", k, aa[k], ab[k], ac[k]);
  // per-atom CSR of the terms each atom belongs to
  std::vector<int> brow(n + 1, 0), bidx(2 * (size_t)nb), arow(n + 1, 0), aidx(3 * (size_t)na);
  for (int k = 0; k < nb; k++) { brow[ba[k] + 1]++; brow[bb[k] + 1]++; }
  for (int k = 0; k < na; k++) { arow[aa[k] + 1]++; arow[ab[k] + 1]++; arow[ac[k] + 1]++; }
  for (int i = 0; i < n; i++) { brow[i + 1] += brow[i]; arow[i + 1] += arow[i]; }
  {
    std::vector<int> fb(brow.begin(), brow.end() - 1), fa(arow.begin(), arow.end() - 1);
    for (int k = 0; k < nb; k++) { bidx[fb[ba[k]]++] = k; bidx[fb[bb[k]]++] = k; }
    for (int k = 0; k < na; k++) { aidx[fa[aa[k]]++] = k; aidx[fa[ab[k]]++] = k; aidx[fa[ac[k]]++] = k; }
  }
  auto vi = [](const int* p, int m) { return std::vector<int>(p, p + m); };
  auto vd = [](const double* p, int m) { return std::vector<double>(p, p + m); };
  if ((r = upload(c, &c->d_b_row, brow))) return r;
  if ((r = upload(c, &c->d_b_idx, bidx))) return r;
  if ((r = upload(c, &c->d_a_row, arow))) return r;
  if ((r = upload(c, &c->d_a_idx, aidx))) return r;
  if ((r = upload(c, &c->d_b_a, vi(ba, nb)))) return r;
  if ((r = upload(c, &c->d_b_b, vi(bb, nb)))) return r;
  if ((r = upload(c, &c->d_b_k, vd(bk, nb)))) return r;
  if ((r = upload(c, &c->d_b_r0, vd(br0, nb)))) return r;
  if ((r = upload(c, &c->d_a_a, vi(aa, na)))) return r;
  if ((r = upload(c, &c->d_a_b, vi(ab, na)))) return r;
  if ((r = upload(c, &c->d_a_c, vi(ac, na)))) return r;
  if ((r = upload(c, &c->d_a_k, vd(ak, na)))) return r;
  if ((r = upload(c, &c->d_a_t0, vd(at0, na)))) return r;
  if ((r = upload(c, &c->d_a_kub, vd(akub, na)))) return r;
  if ((r = upload(c, &c->d_a_rub, vd(arub, na)))) return r;
  c->nbonds = nb;
  c->nangles = na;
  c->h_ba = vi(ba, nb); c->h_bb = vi(bb, nb); c->h_br0 = vd(br0, nb);
  c->h_aa = vi(aa, na); c->h_ab = vi(ab, na); c->h_ac = vi(ac, na); c->h_at0 = vd(at0, na);
  c->cons_built = false;
  c->md_forces_valid = false;
  c->gen++;
  return PM_OK;
}

int pm_set_torsions(pm_context* c, int kind, int nt, const int* atoms, const int* term_row, const double* k,
                    const int* period, const double* phase) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (!c->have_topo) return fail(c, PM_ERR_STATE, "no topology: call pm_set_topology first");
  if (kind != 0 && kind != 1) return fail(c, PM_ERR_ARGUMENT, "pm_set_torsions: kind is 0 (dihedrals) or 1 (impropers)");
  if (nt < 0 || (nt > 0 && (!atoms || !term_row || !k || !period || !phase)))
    return fail(c, PM_ERR_ARGUMENT, "pm_set_torsions: arrays missing");
  const int n = c->n;
  if (nt > 0 && term_row[0] != 0) return fail(c, PM_ERR_ARGUMENT, "pm_set_torsions: term_row[0] must be 0");
  for (int t = 0; t < nt; t++) {
    const int* a = atoms + 4 * (size_t)t;
    for (int j = 0; j < 4; j++)
      if (a[j] < 0 || a[j] >= n) return fail(c, PM_ERR_ARGUMENT, "torsion %d names atom %d of %d", t, a[j], n);
    if (a[0] == a[1] || a[0] == a[2] || a[0] == a[3] || a[1] == a[2] || a[1] == a[3] || a[2] == a[3])
      return fail(c, PM_ERR_ARGUMENT, "torsion %d (%d, %d, %d, %d) repeats an atom", t, a[0], a[1], a[2], a[3]);
    if (term_row[t + 1] <= term_row[t])
      return fail(c, PM_ERR_ARGUMENT, "torsion %d has no term (term_row must increase)", t);
  }
  const int nterm = nt > 0 ? term_row[nt] : 0;
  std::vector<int> row(n + 1, 0), idx(4 * (size_t)nt);
  std::vector<int4> at(nt);
  for (int t = 0; t < nt; t++) {
    at[t] = make_int4(atoms[4 * t], atoms[4 * t + 1], atoms[4 * t + 2], atoms[4 * t + 3]);
    for (int j = 0; j < 4; j++) row[atoms[4 * t + j] + 1]++;
  }
  for (int i = 0; i < n; i++) row[i + 1] += row[i];
  {
    std::vector<int> fill(row.begin(), row.end() - 1);
    for (int t = 0; t < nt; t++)
      for (int j = 0; j < 4; j++) idx[fill[atoms[4 * t + j]]++] = t;
  }
  if ((r = upload(c, &c->d_t_row[kind], row))) return r;
  if ((r = upload(c, &c->d_t_idx[kind], idx))) return r;
  if ((r = upload(c, &c->d_t_atoms[kind], at))) return r;
  if ((r = upload(c, &c->d_t_term[kind], std::vector<int>(term_row, term_row + (nt > 0 ? nt + 1 : 0))))) return r;
  if ((r = upload(c, &c->d_t_k[kind], std::vector<double>(k, k + nterm)))) return r;
  if ((r = upload(c, &c->d_t_n[kind], std::vector<int>(period, period + nterm)))) return r;
  if ((r = upload(c, &c->d_t_phase[kind], std::vector<double>(phase, phase + nterm)))) return r;
  c->ntors[kind] = nt;
  c->gen++;
  c->md_forces_valid = false;
  return PM_OK;
}

int pm_set_boundary(pm_context* c, const double* e1, const double* e2, const double* e3, const double* origin,
                    double cellsize) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (!e1 || !e2 || !e3 || !origin) return fail(c, PM_ERR_ARGUMENT, "pm_set_boundary: null vector");
  if (e1[1] != 0.0 || e1[2] != 0.0 || e2[0] != 0.0 || e2[2] != 0.0 || e3[0] != 0.0 || e3[1] != 0.0)
    return fail(c, PM_ERR_UNSUPPORTED, "[NonbondedFullEwaldSystemForce::initialize] Not orthogonal, aborting.");
  const double lx = e1[0], ly = e2[1], lz = e3[2];
  // PeriodicBoundaryConditions::set (topology/PeriodicBoundaryConditions.cpp:34-77)
  double V = std::fabs(lx * ly * lz);
  if (!(V >= kEpsilon)) return fail(c, PM_ERR_ARGUMENT, "[PeriodicBoundaryConditions::set] No volume, aborting.");
  if (!std::isfinite(V)) return fail(c, PM_ERR_ARGUMENT, "[PeriodicBoundaryConditions::set] Infinite volume, aborting.");
  if (!(cellsize > 0)) return fail(c, PM_ERR_ARGUMENT, "Cubic cellmanager: cutoff > 0 (%s).", fmtnum(cellsize).c_str());
  if (lx < 0 || ly < 0 || lz < 0) return fail(c, PM_ERR_UNSUPPORTED, "cell basis vectors must point along +x, +y, +z");
  PmBox& b = c->box;
  const double L[3] = {lx, ly, lz};
  const double a1x = ly * lz - 0.0 * 0.0, a2y = lz * lx - 0.0 * 0.0, a3z = lx * ly - 0.0 * 0.0;
  b.er[0] = a1x / (lx * a1x + 0.0 * 0.0 + 0.0 * 0.0);
  b.er[1] = a2y / (0.0 * 0.0 + ly * a2y + 0.0 * 0.0);
  b.er[2] = a3z / (0.0 * 0.0 + 0.0 * 0.0 + lz * a3z);
  double mn[3], mx[3];
  for (int k = 0; k < 3; k++) {
    b.L[k] = L[k];
    b.origin[k] = origin[k];
    double a = origin[k] - L[k] * 0.5, bb = origin[k] + L[k] * 0.5;
    mn[k] = std::min(a, bb);
    mx[k] = std::max(a, bb);
    b.h[k] = mx[k] - mn[k];
    b.h2[k] = b.h[k] * 0.5;
    b.delta[k] = origin[k] - mn[k];
    // CubicCellManager::initialize and ArrayCellListStructure::initialize
    double d = mx[k] - mn[k];
    double f = std::floor(d / cellsize);
    c->realcell[k] = d / std::max(1.0, f);
    b.rreal[k] = 1.0 / c->realcell[k];
    b.n[k] = std::max(1, (int)std::floor(d / c->realcell[k] + kEpsilon));
    b.to_fix[k] = 4294967296.0 / L[k];
    b.fix_to_A[k] = L[k] / 4294967296.0;
    double u = mn[k] / L[k];
    u -= std::floor(u);
    b.cell0_fix[k] = u * 4294967296.0;
    b.real_fix[k] = c->realcell[k] * 4294967296.0 / L[k];
  }
  double dx = (lx * 0.5) * (lx * 0.5), dy = (ly * 0.5) * (ly * 0.5), dz = (lz * 0.5) * (lz * 0.5);
  b.d = std::min(dx, std::min(dy, dz));
  long long ncl = (long long)b.n[0] * b.n[1] * b.n[2];
  if (ncl > 64LL * 1024 * 1024)
    return fail(c, PM_ERR_CAPACITY, "Your systems is expanding such that the the ratio simulation box / cell size is to big.");
  c->ncells = (int)ncl;
  c->cellsize = cellsize;
  free_stencils(c);
  c->cells_alloc = 0;
  c->pme_alpha_tables = -1;
  c->have_box = true;
  c->list_stale = true;
  for (auto& f : c->forces) f.cached = false;
  c->md_forces_valid = false;
  c->gen++;
  return PM_OK;
}

int pm_force_create(pm_context* c, const char* id, const double* values, int nvalues, int* handle) {
  if (!c || !handle) return PM_ERR_ARGUMENT;
  if (nvalues < 0 || (nvalues > 0 && !values)) return fail(c, PM_ERR_ARGUMENT, "pm_force_create: values missing");
  ForceDesc f;
  int r = parse_force(c, id, values, nvalues, f);
  if (r) return r;
  c->forces.push_back(f);
  *handle = (int)c->forces.size() - 1;
  return PM_OK;
}

int pm_force_uncache(pm_context* c, int handle) {
  if (!c) return PM_ERR_ARGUMENT;
  if (handle < 0 || handle >= (int)c->forces.size()) return fail(c, PM_ERR_ARGUMENT, "unknown force handle %d", handle);
  c->forces[handle].cached = false;
  c->pme_alpha_tables = -1;
  return PM_OK;
}

static int check_ready(pm_context* c, const int* handles, int nh) {
  if (!c->have_topo) return fail(c, PM_ERR_STATE, "no topology: call pm_set_topology first");
  if (!c->have_box) return fail(c, PM_ERR_STATE, "no boundary conditions: call pm_set_boundary first");
  if (nh <= 0 || !handles) return fail(c, PM_ERR_ARGUMENT, "empty force group");
  for (int i = 0; i < nh; i++)
    if (handles[i] < 0 || handles[i] >= (int)c->forces.size()) return fail(c, PM_ERR_ARGUMENT, "unknown force handle %d", handles[i]);
  return PM_OK;
}

int pm_evaluate(pm_context* c, const int* handles, int nh, const double* positions, double* forces,
                double* energies, int flags) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if ((r = check_ready(c, handles, nh))) return r;
  if (!positions || !forces || !energies) return fail(c, PM_ERR_ARGUMENT, "pm_evaluate: null buffer");
  if (flags & PM_EVAL_MOLVIRIAL)
    return fail(c, PM_ERR_UNSUPPORTED, "pm_evaluate: the molecular virial (ScalarStructure::molecularVirial, "
                                       "OneAtomPair.h:175-183) is not computed on the GPU");
  if (flags & ~(PM_EVAL_VIRIAL | PM_EVAL_MOLVIRIAL))
    return fail(c, PM_ERR_ARGUMENT, "pm_evaluate: unknown flags 0x%x", flags);
  const int n = c->n;
  HIPCHECK(c, hipMemcpyAsync(c->d_pos, positions, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  std::vector<int> hv(handles, handles + nh);
  if ((r = evaluate_group(c, hv, flags))) return r;
  pm_launch_fix_to_double(c->stream, 3 * n, c->d_force, c->d_scratch3);
  std::vector<double> f(3 * (size_t)n), e(PM_NE);
  HIPCHECK(c, hipMemcpyAsync(f.data(), c->d_scratch3, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipMemcpyAsync(e.data(), c->d_energy, PM_NE * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if ((r = check_overflow(c))) return r;
  for (size_t k = 0; k < f.size(); k++) forces[k] += f[k];
  for (int k = 0; k < PM_NE; k++) energies[k] += e[k];
  c->md_forces_valid = false;
  c->gen++;
  return PM_OK;
}

int pm_cell_diagnostics(pm_context* c, const double* positions, double algorithm_cutoff, double pair_cutoff,
                        int* dims, int* cell_of_atom, int* head, int* next, long long* cand, long long* inrc) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (!c->have_topo || !c->have_box) return fail(c, PM_ERR_STATE, "topology and boundary required");
  if (dims) { dims[0] = c->box.n[0]; dims[1] = c->box.n[1]; dims[2] = c->box.n[2]; }
  if (!positions) return PM_OK;
  const int n = c->n, nc = c->ncells;
  if ((r = ensure_cells(c))) return r;
  hipStream_t s = c->stream;
  HIPCHECK(c, hipMemcpyAsync(c->d_pos, positions, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
  c->list_stale = true;   // the binning below replaces the layout's cell lists
  pm_launch_zero_i(s, c->d_bin_count, nc);
  pm_launch_bin(s, c->box, n, c->d_pos, c->d_cell_of, c->d_bin_count, c->d_cell_rank, c->d_key, nullptr);
  pm_launch_scan_cells(s, nc, c->box.n[2], c->d_bin_count, c->d_cell_count, c->d_cell_start, c->d_cell_slot, c->d_counts,
                       c->d_cluster_cell, c->ncl_max, c->m, nullptr, c->d_scan_aux);
  pm_launch_scatter(s, n, c->d_cell_of, c->d_cell_rank, c->d_cell_start, c->d_cell_atoms, nullptr);
  std::vector<int> cell(n), cnt(nc), start(nc), atoms(n);
  HIPCHECK(c, hipMemcpyAsync(cell.data(), c->d_cell_of, n * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipMemcpyAsync(cnt.data(), c->d_cell_count, nc * sizeof(int), hipMemcpyDeviceToHost, s));
  long long* d_cand = nullptr;
  long long* d_inrc = nullptr;
  if (cand || inrc) {
    int4* half;
    int nhalf;
    if ((r = get_stencil(c, algorithm_cutoff, true, &half, &nhalf))) return r;
    if ((r = dalloc(c, &d_cand, nc))) return r;
    if ((r = dalloc(c, &d_inrc, nc))) { dfree(d_cand); return r; }
    pm_launch_diag(s, c->box, n, c->d_pos, c->d_cell_of, c->d_cell_start, c->d_cell_atoms, half, nhalf, c->d_mol,
                   c->d_ex_row, c->d_ex_partner, c->d_ex_cls, pair_cutoff * pair_cutoff, nc, c->d_cell_count,
                   d_cand, d_inrc);
    std::vector<long long> hc(nc), hi(nc);
    HIPCHECK(c, hipMemcpyAsync(hc.data(), d_cand, nc * sizeof(long long), hipMemcpyDeviceToHost, s));
    HIPCHECK(c, hipMemcpyAsync(hi.data(), d_inrc, nc * sizeof(long long), hipMemcpyDeviceToHost, s));
    HIPCHECK(c, hipStreamSynchronize(s));
    if (cand) std::copy(hc.begin(), hc.end(), cand);
    if (inrc) std::copy(hi.begin(), hi.end(), inrc);
    dfree(d_cand);
    dfree(d_inrc);
  }
  HIPCHECK(c, hipStreamSynchronize(s));
  if (cell_of_atom) std::copy(cell.begin(), cell.end(), cell_of_atom);
  if (head || next) {
    // linked lists in the reference's order: atoms prepended from N-1 down to 0
    // (buildCellLists.cpp:19-36), i.e. ascending index within each cell
    std::vector<int> hd(nc, -1), nx(n, -1);
    for (int i = n - 1; i >= 0; i--) { nx[i] = hd[cell[i]]; hd[cell[i]] = i; }
    if (head) std::copy(hd.begin(), hd.end(), head);
    if (next) std::copy(nx.begin(), nx.end(), next);
  }
  return PM_OK;
}

int pm_md_setup(pm_context* c, const int* handles, int nh, int integrator, double dt, double temperature,
                double gamma, unsigned long long seed) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = check_ready(c, handles, nh))) return r;
  if (integrator != PM_INTEGRATOR_LEAPFROG && integrator != PM_INTEGRATOR_LANGEVIN)
    return fail(c, PM_ERR_ARGUMENT, "unknown integrator %d", integrator);
  if (!(dt > 0)) return fail(c, PM_ERR_ARGUMENT, "timestep must be > 0");
  if (integrator == PM_INTEGRATOR_LANGEVIN && !(gamma > 0))
    return fail(c, PM_ERR_ARGUMENT, "Langevin gamma must be > 0");
  c->md_handles.assign(handles, handles + nh);
  c->md_integrator = integrator;
  c->md_dt = dt;
  c->md_temp = temperature;
  c->md_gamma = gamma;
  c->md_seed = seed;
  c->md_counter = 0;
  c->md_forces_valid = false;
  c->gen++;
  if (!c->d_md && (r = dalloc(c, &c->d_md, 1))) return r;
  PmMdDevice md{0ull, temperature};
  HIPCHECK(c, hipMemcpy(c->d_md, &md, sizeof(md), hipMemcpyHostToDevice));
  return PM_OK;
}

int pm_md_set_state(pm_context* c, const double* x, const double* v) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (!c->have_topo) return fail(c, PM_ERR_STATE, "no topology");
  const size_t bytes = 3 * (size_t)c->n * sizeof(double);
  if (x) HIPCHECK(c, hipMemcpy(c->d_pos, x, bytes, hipMemcpyHostToDevice));
  if (v) HIPCHECK(c, hipMemcpy(c->d_vel, v, bytes, hipMemcpyHostToDevice));
  // SHAKE's reference positions start at the new positions
  if (x && c->d_shake_last) HIPCHECK(c, hipMemcpy(c->d_shake_last, x, bytes, hipMemcpyHostToDevice));
  // the forces follow the positions only
  if (x) c->md_forces_valid = false;
  return PM_OK;
}

int pm_md_set_constraints(pm_context* c, double shake_eps, int shake_max_iter, double rattle_eps,
                          int rattle_max_iter) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (!c->have_topo) return fail(c, PM_ERR_STATE, "no topology: call pm_set_topology first");
  if (shake_max_iter < 0 || rattle_max_iter < 0 || shake_max_iter > 100000 || rattle_max_iter > 100000 ||
      !(shake_eps == shake_eps) || !(rattle_eps == rattle_eps))
    return fail(c, PM_ERR_ARGUMENT, "constraint settings: epsilon must be a number, max iterations 0..100000");
  // a modifier is attached only with epsilon > 0 and maxIter > 0 (protomol.cpp:553-578)
  const bool shake = shake_eps > 0.0 && shake_max_iter > 0, rattle = rattle_eps > 0.0 && rattle_max_iter > 0;
  c->shake_eps = shake ? shake_eps : 0.0;
  c->shake_maxit = shake ? shake_max_iter : 0;
  c->rattle_eps = rattle ? rattle_eps : 0.0;
  c->rattle_maxit = rattle ? rattle_max_iter : 0;
  if (shake || rattle) {
    if ((r = build_constraints(c))) return r;
    if ((r = ensure_constraint_err(c, std::max(c->shake_maxit, c->rattle_maxit)))) return r;
    // myLastPositions = the positions when the modifier is initialized
    // (ModifierMetaRattleShake.cpp:17)
    if (c->d_pos) HIPCHECK(c, hipMemcpy(c->d_shake_last, c->d_pos, 3 * (size_t)c->n * sizeof(double),
                                        hipMemcpyDeviceToDevice));
  }
  c->gen++;   // captured steps are stale
  return PM_OK;
}

int pm_md_constraint_count(pm_context* c, int* count) {
  if (!c || !count) return PM_ERR_ARGUMENT;
  int r;
  if (!c->have_topo) return fail(c, PM_ERR_STATE, "no topology: call pm_set_topology first");
  if ((r = set_device(c))) return r;
  if ((r = build_constraints(c))) return r;
  *count = c->cons_n;
  return PM_OK;
}

// one SHAKE (rattle 0) or RATTLE (rattle 1) application on host arrays
static int constraint_call(pm_context* c, int rattle, const double* x_in, double* x_out, double* v,
                           const double* x_last, double dt_fs, double eps, int max_iter, int* sweeps) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (!c->have_topo) return fail(c, PM_ERR_STATE, "no topology: call pm_set_topology first");
  if (!x_in || !v || (!rattle && (!x_last || !x_out)))
    return fail(c, PM_ERR_ARGUMENT, "constraints: null positions, velocities or reference positions");
  if (!(dt_fs > 0.0) || !(eps > 0.0) || max_iter <= 0 || max_iter > 100000)
    return fail(c, PM_ERR_ARGUMENT, "constraints: timestep and epsilon must be > 0, max iterations 1..100000");
  if ((r = build_constraints(c))) return r;
  if ((r = ensure_constraint_err(c, max_iter))) return r;
  const size_t n3 = 3 * (size_t)c->n, bytes = n3 * sizeof(double);
  if (!c->d_cons_x) {
    if ((r = dalloc(c, &c->d_cons_x, n3))) return r;
    if ((r = dalloc(c, &c->d_cons_v, n3))) return r;
    if ((r = dalloc(c, &c->d_cons_xl, n3))) return r;
  }
  hipStream_t s = c->stream;
  HIPCHECK(c, hipMemcpyAsync(c->d_cons_x, x_in, bytes, hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipMemcpyAsync(c->d_cons_v, v, bytes, hipMemcpyHostToDevice, s));
  if (!rattle) HIPCHECK(c, hipMemcpyAsync(c->d_cons_xl, x_last, bytes, hipMemcpyHostToDevice, s));
  int sw = 0;
  if (c->cons_ngroup > 0) {
    PmConstraints k = constraint_launch(c, eps, max_iter, dt_fs);
    pm_launch_constraints(s, k, rattle, c->d_cons_x, c->d_cons_v, c->d_cons_xl);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipMemcpyAsync(&sw, c->d_cons_sweeps, sizeof(int), hipMemcpyDeviceToHost, s));
  }
  if (!rattle) HIPCHECK(c, hipMemcpyAsync(x_out, c->d_cons_x, bytes, hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipMemcpyAsync(v, c->d_cons_v, bytes, hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipStreamSynchronize(s));
  if (sweeps) *sweeps = sw;
  return PM_OK;
}

int pm_shake(pm_context* c, double* x, double* v, const double* x_last, double dt_fs, double eps, int max_iter,
             int* sweeps) {
  return constraint_call(c, 0, x, x, v, x_last, dt_fs, eps, max_iter, sweeps);
}

int pm_rattle(pm_context* c, const double* x, double* v, double dt_fs, double eps, int max_iter, int* sweeps) {
  return constraint_call(c, 1, x, nullptr, v, nullptr, dt_fs, eps, max_iter, sweeps);
}

int pm_shake_noalloc(pm_context* c, double* x, double* v, const double* x_last, double dt_fs, double eps,
                     int max_iter, int* sweeps) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (!c->have_topo) return fail(c, PM_ERR_STATE, "no topology: call pm_set_topology first");
  if (!x || !v || !x_last)
    return fail(c, PM_ERR_ARGUMENT, "pm_shake_noalloc: null positions, velocities or reference positions");
  if (!(dt_fs > 0.0) || !(eps > 0.0) || max_iter <= 0 || max_iter > 100000)
    return fail(c, PM_ERR_ARGUMENT, "pm_shake_noalloc: timestep and epsilon must be > 0, max iterations 1..100000");
  // the constraint list is set-up data (pm_md_set_constraints or an earlier
  // pm_shake / pm_rattle built it); building it here would allocate
  if (!c->cons_built)
    return fail(c, PM_ERR_STATE, "pm_shake_noalloc: no constraint list: call pm_md_set_constraints (or pm_shake) "
                                 "once first");
  const size_t n3 = 3 * (size_t)c->n;
  const size_t nwave = 4 * (size_t)std::max((c->cons_ngroup + 255) / 256, 1);
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t vec = up(n3 * sizeof(double));
  const size_t err = up((size_t)(max_iter + 1) * nwave * sizeof(double));
  const size_t need = 4 * vec + err + 256;
  if (need > PM_NOALLOC_SCRATCH_BYTES)
    return fail(c, PM_ERR_ARGUMENT,
                "pm_shake_noalloc: needs %zu bytes of work space (%d atoms, %d sweeps) but the library's fixed "
                "work space holds %u bytes; use pm_shake", need, c->n, max_iter + 1, PM_NOALLOC_SCRATCH_BYTES);
  unsigned char* base = static_cast<unsigned char*>(pm_scratch_address(c->device));
  if (!base) return fail(c, PM_ERR_STATE, "pm_shake_noalloc: the work space of this device was not set up");
  double* dx = reinterpret_cast<double*>(base);
  double* dv = reinterpret_cast<double*>(base + vec);
  double* dxl = reinterpret_cast<double*>(base + 2 * vec);
  double* dscr = reinterpret_cast<double*>(base + 3 * vec);
  double* derr = reinterpret_cast<double*>(base + 4 * vec);
  int* dsweeps = reinterpret_cast<int*>(base + 4 * vec + err);
  const size_t bytes = n3 * sizeof(double);
  hipStream_t s = c->stream;
  HIPCHECK(c, hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipMemcpyAsync(dv, v, bytes, hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipMemcpyAsync(dxl, x_last, bytes, hipMemcpyHostToDevice, s));
  int sw = 0;
  if (c->cons_ngroup > 0) {
    // every wave of the probe writes its error part for every sweep count and
    // the apply kernel writes the sweep count, so nothing here starts at zero
    PmConstraints k = constraint_launch(c, eps, max_iter, dt_fs);
    k.err = derr;
    k.sweeps = dsweeps;
    k.scratch = dscr;
    pm_launch_constraints(s, k, 0, dx, dv, dxl);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipMemcpyAsync(&sw, dsweeps, sizeof(int), hipMemcpyDeviceToHost, s));
  }
  HIPCHECK(c, hipMemcpyAsync(x, dx, bytes, hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipMemcpyAsync(v, dv, bytes, hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipStreamSynchronize(s));
  if (sweeps) *sweeps = sw;
  return PM_OK;
}

}  // extern "C"

namespace {

// One MD step without its closing half kick, as queued on the context's
// streams: kind 0 = a middle leapfrog step (LeapfrogIntegrator::doKickdoDrift,
// then the forces), kind 1 = a whole Langevin step (StandardIntegrator::run:
// half kick, drift, forces, half kick).
// SHAKE / RATTLE in MD steps (pm_md_set_constraints)
bool md_shake(const pm_context* c) { return c->shake_maxit > 0 && c->cons_built && c->cons_ngroup > 0; }
bool md_rattle(const pm_context* c) { return c->rattle_maxit > 0 && c->cons_built && c->cons_ngroup > 0; }
void md_shake_step(pm_context* c) {
  PmConstraints k = constraint_launch(c, c->shake_eps, c->shake_maxit, c->md_dt);
  k.sweeps = nullptr;
  pm_launch_constraints(c->stream, k, 0, c->d_pos, c->d_vel, c->d_shake_last);
}
void md_rattle_step(pm_context* c) {
  PmConstraints k = constraint_launch(c, c->rattle_eps, c->rattle_maxit, c->md_dt);
  k.sweeps = nullptr;
  pm_launch_constraints(c->stream, k, 1, c->d_pos, c->d_vel, c->d_shake_last);
}

// An evaluation after a position update with PmStepWork failed: the forces
// were cleared and the binning counts may hold a binning that no scan
// consumed; the next run starts from a fresh force evaluation and zero counts.
void md_step_failed(pm_context* c) {
  c->md_forces_valid = false;
  c->md_prebinned = false;
  c->disp_prechecked = false;
  if (c->d_bin_count && c->ncells > 0) pm_launch_zero_i(c->stream, c->d_bin_count, c->ncells);
}

int md_unit(pm_context* c, int kind, bool energy) {
  hipStream_t s = c->stream;
  const int n = c->n;
  const double h = c->md_dt * kInvTimeFactor;
  int r;
  PmStepWork w{};
  const bool shake = md_shake(c), rattle = md_rattle(c);
  if (kind == 0) {
    {
      TimedScope ts(c, PM_T_INTEGRATE);
      if (rattle) {
        // LeapfrogIntegrator::doKickdoDrift with a post-step modifier
        // (LeapfrogIntegrator.cpp:72-84): half kick, RATTLE, half kick, drift
        // (StandardIntegrator::doHalfKick, STSIntegrator::doDrift), SHAKE
        const double hh = 0.5 * c->md_dt * kInvTimeFactor;
        pm_launch_kick(s, n, 2, hh, c->d_mass, c->d_force, c->d_vel, c->d_pos);
        md_rattle_step(c);
        pm_launch_kick(s, n, 2, hh, c->d_mass, c->d_force, c->d_vel, c->d_pos);
        pm_launch_drift(s, n, h, c->d_vel, c->d_pos);
      } else {
        // the displacement test and binning need the final positions: not
        // in the update when SHAKE moves them afterwards
        const bool sw = !shake && md_step_work(c, &w);
        pm_launch_kick(s, n, 1, h, c->d_mass, c->d_force, c->d_vel, c->d_pos, sw ? &w : nullptr, sw ? 1 : 0);
      }
      if (shake) md_shake_step(c);
    }
    if ((r = evaluate_group(c, c->md_handles, 0, energy))) md_step_failed(c);
    return r;
  }
  const double g = c->md_gamma / (1000 * kInvTimeFactor);
  {
    TimedScope ts(c, PM_T_INTEGRATE);
    const bool sw = !shake && md_step_work(c, &w);
    pm_launch_langevin_kick(s, n, h, g, c->d_mass, c->d_force, c->d_vel, c->md_seed, c->d_md, 0, sw ? 1 : 0);
    pm_launch_drift(s, n, h, c->d_vel, c->d_pos, sw ? &w : nullptr);
    if (shake) md_shake_step(c);   // postDriftOrNextModify (STSIntegrator.cpp:42-47)
  }
  if ((r = evaluate_group(c, c->md_handles, 0, energy))) {
    md_step_failed(c);
    return r;
  }
  TimedScope ts(c, PM_T_INTEGRATE);
  pm_launch_langevin_kick(s, n, h, g, c->d_mass, c->d_force, c->d_vel, c->md_seed, c->d_md, 1);
  if (rattle) md_rattle_step(c);   // postStepModify (StandardIntegrator.cpp:28-34)
  pm_launch_md_advance(s, c->d_md);
  c->md_counter += 2;
  return PM_OK;
}

void drop_graph(pm_context* c) {
  if (c->md_exec) hipGraphExecDestroy(c->md_exec);
  c->md_exec = nullptr;
  c->md_exec_steps = 0;
  c->md_exec_kind = -1;
}

// `count` units of `kind` without energies: replays of a captured graph of up
// to 8 units where allowed (no per-kernel timing requested), the rest queued
// directly.  The graph holds no memset or copy nodes: only kernels and the
// fork/join of the second stream.
int md_units(pm_context* c, int kind, int count) {
  int r;
  // graphs always hold 8 steps: a run of fewer steps captures the graph the
  // next runs replay and launches its own steps directly, so that runs of
  // different lengths never capture again
  const int G = 8;
  const bool use_graph = c->graph_option != 0 && c->timing_mask == 0 && count >= 2;
  int done = 0;
  if (use_graph) {
    if (!c->md_exec || c->md_exec_gen != c->gen || c->md_exec_kind != kind || c->md_exec_steps != G) {
      drop_graph(c);
      // one direct unit first, so that every buffer and table exists before capture
      if ((r = md_unit(c, kind, false))) return r;
      done = 1;
      const unsigned long long gen0 = c->gen;
      hipGraph_t g = nullptr;
      HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = PM_OK;
      for (int u = 0; u < G && rc == PM_OK; u++) rc = md_unit(c, kind, false);
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (rc != PM_OK) { if (g) hipGraphDestroy(g); return rc; }
      if (e != hipSuccess) return fail(c, PM_ERR_DEVICE, "HIP error %s ending the capture of an MD step", hipGetErrorString(e));
      // the capture queued nothing: the units it recorded run when the graph is launched
      c->md_counter -= (kind == 1) ? 2ull * G : 0ull;
      e = hipGraphInstantiate(&c->md_exec, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) { c->md_exec = nullptr; return fail(c, PM_ERR_DEVICE, "HIP error %s instantiating the MD step graph", hipGetErrorString(e)); }
      if (c->gen != gen0) { drop_graph(c); return fail(c, PM_ERR_STATE, "internal: buffers changed while an MD step was captured"); }
      c->md_exec_gen = c->gen;
      c->md_exec_kind = kind;
      c->md_exec_steps = G;
    }
    while (count - done >= G) {
      HIPCHECK(c, hipGraphLaunch(c->md_exec, c->stream));
      if (kind == 1) c->md_counter += 2ull * G;
      done += G;
    }
  }
  for (; done < count; done++)
    if ((r = md_unit(c, kind, false))) return r;
  return PM_OK;
}

}  // namespace

extern "C" {

int pm_md_run(pm_context* c, int nsteps) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (c->md_handles.empty()) return fail(c, PM_ERR_STATE, "pm_md_setup not called");
  if (nsteps < 1) return PM_OK;
  hipStream_t s = c->stream;
  const int n = c->n;
  const double h = c->md_dt * kInvTimeFactor;
  if (!c->md_forces_valid) {
    // initializeForces (StandardIntegrator.cpp:46-59)
    if ((r = evaluate_group(c, c->md_handles, 0, false))) return r;
    c->md_forces_valid = true;
  }
  if (c->md_integrator == PM_INTEGRATOR_LEAPFROG) {
    // LeapfrogIntegrator::run (integrators/LeapfrogIntegrator.cpp:107-121):
    // half kick + drift, forces, (kick + drift, forces) x (n - 1), half kick;
    // energies of the last force evaluation only
    {
      TimedScope ts(c, PM_T_INTEGRATE);
      PmStepWork w{};
      const bool shake = md_shake(c);
      const bool sw = !shake && md_step_work(c, &w);
      pm_launch_kick(s, n, 0, h, c->d_mass, c->d_force, c->d_vel, c->d_pos, sw ? &w : nullptr, sw ? 1 : 0);
      if (shake) md_shake_step(c);   // postDriftOrNextModify (LeapfrogIntegrator.cpp:66)
    }
    if ((r = evaluate_group(c, c->md_handles, 0, nsteps == 1))) {
      md_step_failed(c);
      return r;
    }
    if (nsteps >= 3 && (r = md_units(c, 0, nsteps - 2))) return r;
    if (nsteps >= 2 && (r = md_unit(c, 0, true))) return r;
    pm_launch_kick(s, n, 2, 0.5 * c->md_dt * kInvTimeFactor, c->d_mass, c->d_force, c->d_vel, c->d_pos);
    if (md_rattle(c)) md_rattle_step(c);   // postStepModify (LeapfrogIntegrator.cpp:119-120)
  } else {
    // StandardIntegrator::run with the Langevin half kick
    // (StandardIntegrator.cpp:24-35, LangevinLeapfrogIntegrator.cpp:50-79)
    if (nsteps >= 2 && (r = md_units(c, 1, nsteps - 1))) return r;
    if ((r = md_unit(c, 1, true))) return r;
  }
  HIPCHECK(c, hipGetLastError());
  return PM_OK;
}

int pm_md_autotune(pm_context* c, int trial_steps, int* cluster_size, int* list_skin, double* ms_per_step) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (c->md_handles.empty()) return fail(c, PM_ERR_STATE, "pm_md_setup not called");
  if (trial_steps < 4 || trial_steps > 100000)
    return fail(c, PM_ERR_ARGUMENT, "pm_md_autotune: trial_steps must lie in [4, 100000] (got %d)", trial_steps);
  if (!c->d_pos || !c->d_vel) return fail(c, PM_ERR_STATE, "pm_md_set_state not called");
  hipStream_t s = c->stream;
  const size_t n3 = 3 * (size_t)c->n;
  // the state every trial starts from, restored at the end
  std::vector<double> x(n3), v(n3);
  PmMdDevice md{};
  HIPCHECK(c, hipStreamSynchronize(s));
  HIPCHECK(c, hipMemcpy(x.data(), c->d_pos, n3 * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(v.data(), c->d_vel, n3 * sizeof(double), hipMemcpyDeviceToHost));
  if (c->d_md) HIPCHECK(c, hipMemcpy(&md, c->d_md, sizeof(md), hipMemcpyDeviceToHost));
  const unsigned long long counter0 = c->md_counter;
  const int m0 = c->m_override, skin0 = c->skin_milli;
  auto restore = [&]() -> int {
    HIPCHECK(c, hipMemcpy(c->d_pos, x.data(), n3 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_vel, v.data(), n3 * sizeof(double), hipMemcpyHostToDevice));
    if (c->d_md) HIPCHECK(c, hipMemcpy(c->d_md, &md, sizeof(md), hipMemcpyHostToDevice));
    c->md_counter = counter0;
    c->md_forces_valid = false;
    return PM_OK;
  };
  auto configure = [&](int m, int skin) {
    if (c->m_override != m) c->gen++;
    c->m_override = m;
    c->skin_milli = skin;
    c->list_stale = true;
    c->md_forces_valid = false;
    c->gen++;
  };
  const int ms_opts[2] = {8, 4};
  const int skin_opts[4] = {0, 400, 800, 1200};
  double best = 1e300;
  int best_m = 8, best_skin = 600;
  for (int mi = 0; mi < 2; mi++)
    for (int si = 0; si < 4; si++) {
      configure(ms_opts[mi], skin_opts[si]);
      if ((r = restore())) break;
      // warm-up: allocations, graph capture, the first list build
      if ((r = pm_md_run(c, std::max(3, trial_steps / 4)))) break;
      HIPCHECK(c, hipStreamSynchronize(s));
      const auto t0 = std::chrono::steady_clock::now();
      if ((r = pm_md_run(c, trial_steps))) break;
      HIPCHECK(c, hipStreamSynchronize(s));
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / trial_steps;
      if (ms < best) { best = ms; best_m = ms_opts[mi]; best_skin = skin_opts[si]; }
    }
  if (r) {
    configure(m0, skin0);
    restore();
    return r;
  }
  configure(best_m, best_skin);
  if ((r = restore())) return r;
  if (cluster_size) *cluster_size = best_m;
  if (list_skin) *list_skin = best_skin;
  if (ms_per_step) *ms_per_step = best;
  return PM_OK;
}

int pm_md_get_state(pm_context* c, double* x, double* v, double* energies) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  const size_t bytes = 3 * (size_t)c->n * sizeof(double);
  if (x) HIPCHECK(c, hipMemcpyAsync(x, c->d_pos, bytes, hipMemcpyDeviceToHost, c->stream));
  if (v) HIPCHECK(c, hipMemcpyAsync(v, c->d_vel, bytes, hipMemcpyDeviceToHost, c->stream));
  if (energies) HIPCHECK(c, hipMemcpyAsync(energies, c->d_energy, PM_NE * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return check_overflow(c);
}

int pm_last_stats(pm_context* c, long long* stats, int nstats) {
  if (!c || !stats) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if ((r = list_stats(c))) return r;
  for (int k = 0; k < nstats && k < 6; k++) stats[k] = c->last_stats[k];
  return PM_OK;
}

int pm_set_timing(pm_context* c, unsigned mask) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  harvest_timing(c);
  c->timing_mask = mask;
  for (int k = 0; k < PM_TIMING_CATEGORIES; k++) { c->t_ms[k] = 0; c->t_count[k] = 0; }
  return PM_OK;
}

int pm_get_timing(pm_context* c, double* ms, long long* count, int ncat) {
  if (!c || !ms || !count) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream2));
  harvest_timing(c);
  for (int k = 0; k < ncat && k < PM_TIMING_CATEGORIES; k++) { ms[k] = c->t_ms[k]; count[k] = c->t_count[k]; }
  return PM_OK;
}

// the adjoint work buffers for the current atom and slot counts
static int ensure_adjoint(pm_context* c) {
  const long long ns = (long long)c->nslot_max + 1;
  int r;
  if (c->adj_nslot < ns) {
    if ((r = dalloc(c, &c->d_adj_sv, (size_t)ns)) || (r = dalloc(c, &c->d_adj_slot, 3 * (size_t)ns))) return r;
    c->adj_nslot = ns;
  }
  if (c->adj_n < c->n) {
    const size_t n3 = 3 * (size_t)c->n;
    if ((r = dalloc(c, &c->d_adj_fix, n3)) || (r = dalloc(c, &c->d_adj_v, n3)) || (r = dalloc(c, &c->d_adj_out, n3)))
      return r;
    c->adj_n = c->n;
  }
  return PM_OK;
}

static int adjoint_check(pm_context* c, const int* handles, int nh) {
  int r;
  if ((r = check_ready(c, handles, nh))) return r;
  for (int k = 0; k < nh; k++)
    if (c->forces[handles[k]].bonded)
      return fail(c, PM_ERR_UNSUPPORTED, "Hessian-vector products and parameter gradients cover the nonbonded cutoff and PME forces, not '%s'", c->forces[handles[k]].id.c_str());
  if (choose_m(c) != 8) return fail(c, PM_ERR_UNSUPPORTED, "the adjoint kernels use clusters of 8 atoms (PM_OPT_CLUSTER_SIZE 4 is set)");
  if ((r = ensure_cells(c))) return r;
  return ensure_adjoint(c);
}

// one adjoint pass over the positions already in d_pos: mode 1 H v with the
// direction in d_adj_v, mode 2 parameter gradients; results in d_adj_fix
// (fixed point) and d_adj_out (FP64 parts)
static int adjoint_pass(pm_context* c, const int* handles, int nh, int mode) {
  AdjointSpec a;
  a.mode = mode;
  a.slot_v = c->d_adj_sv; a.slot = c->d_adj_slot; a.fix = c->d_adj_fix; a.v = c->d_adj_v;
  if (mode == 1) a.hv = c->d_adj_out; else a.dq = c->d_adj_out;
  std::vector<int> hs(handles, handles + nh);
  return evaluate_group(c, hs, 0, false, &a);
}

static int adjoint_run(pm_context* c, const int* handles, int nh, const double* positions, const double* v,
                       int mode, std::vector<long long>& fix, std::vector<double>& dbl) {
  int r;
  if (!positions) return fail(c, PM_ERR_ARGUMENT, "positions required");
  if ((r = adjoint_check(c, handles, nh))) return r;
  const int n = c->n;
  hipStream_t s = c->stream;
  const size_t bytes = 3 * (size_t)n * sizeof(double);
  hipError_t e = hipMemcpyAsync(c->d_pos, positions, bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(c->d_adj_v, v ? v : positions, bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return fail(c, PM_ERR_DEVICE, "HIP error %s copying the inputs", hipGetErrorString(e));
  c->md_forces_valid = false;
  if ((r = adjoint_pass(c, handles, nh, mode))) { hipStreamSynchronize(s); return r; }
  fix.assign(3 * (size_t)n, 0);
  dbl.assign(mode == 1 ? 3 * (size_t)n : (size_t)n, 0.0);
  e = hipMemcpyAsync(fix.data(), c->d_adj_fix, fix.size() * sizeof(long long), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dbl.data(), c->d_adj_out, dbl.size() * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(c, PM_ERR_DEVICE, "HIP error %s copying the results", hipGetErrorString(e));
  return check_overflow(c);
}

int pm_hessian_vec(pm_context* c, const int* handles, int nh, const double* positions, const double* v,
                   double* hv) {
  if (!c || !v || !hv) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  std::vector<long long> fix;
  std::vector<double> dbl;
  if ((r = adjoint_run(c, handles, nh, positions, v, 1, fix, dbl))) return r;
  for (size_t k = 0; k < fix.size(); k++) hv[k] = (double)fix[k] * PM_FORCE_UNFIX + dbl[k];
  return PM_OK;
}

int pm_param_gradients(pm_context* c, const int* handles, int nh, const double* positions, double* dE_dq,
                       double* dE_dsigma, double* dE_depsilon) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if ((dE_dsigma || dE_depsilon) && c->lj_R.empty())
    return fail(c, PM_ERR_STATE, "gradients with respect to sigma and epsilon need per-atom LJ parameters (pm_set_lj_per_atom)");
  std::vector<long long> fix;
  std::vector<double> dbl;
  if ((r = adjoint_run(c, handles, nh, positions, nullptr, 2, fix, dbl))) return r;
  const double rfac = std::pow(2.0, -5.0 / 6.0);   // dR/dsigma
  for (int i = 0; i < c->n; i++) {
    if (dE_dq) dE_dq[i] = (double)fix[3 * i] * PM_FORCE_UNFIX + dbl[i];
    if (dE_dsigma) dE_dsigma[i] = (double)fix[3 * i + 1] * PM_FORCE_UNFIX * rfac;
    // de/deps = 1/(2 sqrt(eps)): no finite derivative at eps = 0 (reported 0)
    if (dE_depsilon) dE_depsilon[i] = c->lj_e[i] > 0 ? (double)fix[3 * i + 2] * PM_FORCE_UNFIX / (2.0 * c->lj_e[i]) : 0.0;
  }
  return PM_OK;
}

int pm_set_charges(pm_context* c, const double* q) {
  if (!c || !q) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (!c->have_topo) return fail(c, PM_ERR_STATE, "no topology: call pm_set_topology first");
  for (int i = 0; i < c->n; i++)
    if (!std::isfinite(q[i])) return fail(c, PM_ERR_ARGUMENT, "charge %d is not finite", i);
  c->q.assign(q, q + c->n);
  // same size: the copy keeps the device address, so captured MD steps stay valid
  HIPCHECK(c, hipMemcpy(c->d_q, c->q.data(), c->q.size() * sizeof(double), hipMemcpyHostToDevice));
  c->md_forces_valid = false;
  return PM_OK;
}

// Eigenvalues and eigenvectors of a symmetric m x m matrix a (row major,
// destroyed) by cyclic Jacobi rotations (Golub & Van Loan, Matrix
// Computations, 4th ed., Algorithm 8.5.3): z column j holds the eigenvector
// of d[j].
static void jacobi_eigen(int m, std::vector<double>& a, std::vector<double>& d, std::vector<double>& z) {
  z.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; i++) z[(size_t)i * m + i] = 1.0;
  auto A = [&](int i, int j) -> double& { return a[(size_t)i * m + j]; };
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++) {
        tot += A(i, j) * A(i, j);
        if (i != j) off += A(i, j) * A(i, j);
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < m - 1; p++)
      for (int q = p + 1; q < m; q++) {
        const double apq = A(p, q);
        if (apq == 0.0) continue;
        const double tau = (A(q, q) - A(p, p)) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
        for (int k = 0; k < m; k++) {   // A <- A J
          const double akp = A(k, p), akq = A(k, q);
          A(k, p) = cs * akp - sn * akq;
          A(k, q) = sn * akp + cs * akq;
        }
        for (int k = 0; k < m; k++) {   // A <- J^T A
          const double apk = A(p, k), aqk = A(q, k);
          A(p, k) = cs * apk - sn * aqk;
          A(q, k) = sn * apk + cs * aqk;
        }
        for (int k = 0; k < m; k++) {   // Z <- Z J
          const double zkp = z[(size_t)k * m + p], zkq = z[(size_t)k * m + q];
          z[(size_t)k * m + p] = cs * zkp - sn * zkq;
          z[(size_t)k * m + q] = sn * zkp + cs * zkq;
        }
      }
  }
  d.resize(m);
  for (int i = 0; i < m; i++) d[i] = A(i, i);
}

int pm_lanczos(pm_context* c, const int* handles, int nh, const double* positions, const double* v0, int steps,
               int mass_weighted, int nvec, double* ritz_values, double* ritz_vectors, double* residuals,
               double* alpha_beta) {
  if (!c || !v0 || !ritz_values) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  if (!positions) return fail(c, PM_ERR_ARGUMENT, "positions required");
  const int n = c->n, n3 = 3 * n;
  if (steps < 1 || steps > PM_LANCZOS_MAX_STEPS || steps > n3)
    return fail(c, PM_ERR_ARGUMENT, "pm_lanczos: steps must lie in [1, min(%d, 3n)] (got %d)", PM_LANCZOS_MAX_STEPS, steps);
  if (nvec < 0 || nvec > steps) return fail(c, PM_ERR_ARGUMENT, "pm_lanczos: nvec must lie in [0, steps] (got %d)", nvec);
  if (nvec > 0 && !ritz_vectors) return fail(c, PM_ERR_ARGUMENT, "pm_lanczos: ritz_vectors required for nvec > 0");
  double nrm0 = 0.0;
  for (int k = 0; k < n3; k++) nrm0 += v0[k] * v0[k];
  if (!(nrm0 > 0.0) || !std::isfinite(nrm0)) return fail(c, PM_ERR_ARGUMENT, "pm_lanczos: the start vector must be finite and nonzero");
  if (mass_weighted)
    for (int i = 0; i < n; i++)
      if (!(c->mass[i] > 0.0)) return fail(c, PM_ERR_ARGUMENT, "pm_lanczos: mass weighting needs positive masses (atom %d)", i);
  if ((r = adjoint_check(c, handles, nh))) return r;
  hipStream_t s = c->stream;
  const int nblk = pm_krylov_dot_blocks(n3);
  const int m = steps;
  // device work space: basis V [m + 1][n3], w, 1/sqrt(mass), partial sums,
  // coefficients [2][m + 1], alpha [m], beta [m], tridiagonal eigenvectors
  double *V = nullptr, *w = nullptr, *wt = nullptr, *part = nullptr, *coef = nullptr, *ab = nullptr, *S = nullptr,
         *Y = nullptr;
  auto cleanup = [&]() { dfree(V); dfree(w); dfree(wt); dfree(part); dfree(coef); dfree(ab); dfree(S); dfree(Y); };
  if ((r = dalloc(c, &V, (size_t)(m + 1) * n3)) || (r = dalloc(c, &w, (size_t)n3)) ||
      (r = dalloc(c, &wt, (size_t)n)) || (r = dalloc(c, &part, (size_t)(m + 1) * nblk)) ||
      (r = dalloc(c, &coef, 2 * (size_t)(m + 1))) || (r = dalloc(c, &ab, 2 * (size_t)m)) ||
      (r = dalloc(c, &S, (size_t)std::max(1, nvec) * m)) || (r = dalloc(c, &Y, nvec > 0 ? (size_t)nvec * n3 : 1))) {
    cleanup();
    return r;
  }
  std::vector<double> h(n3);
  const double inv0 = 1.0 / std::sqrt(nrm0);
  for (int k = 0; k < n3; k++) h[k] = v0[k] * inv0;
  std::vector<double> minv(n);
  for (int i = 0; i < n; i++) minv[i] = 1.0 / std::sqrt(c->mass[i]);
  const size_t bytes = (size_t)n3 * sizeof(double);
  hipError_t e = hipMemcpyAsync(c->d_pos, positions, bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(V, h.data(), bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(wt, minv.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) { hipStreamSynchronize(s); cleanup(); return fail(c, PM_ERR_DEVICE, "HIP error %s copying the inputs", hipGetErrorString(e)); }
  c->md_forces_valid = false;
  const double* mw = mass_weighted ? wt : nullptr;
  double* alpha = ab;
  double* beta = ab + m;
  for (int j = 0; j < m; j++) {
    double* vj = V + (size_t)j * n3;
    // w = M^-1/2 H M^-1/2 v_j (M = 1 without mass weighting)
    pm_launch_scale_copy(s, n3, vj, mw, c->d_adj_v);
    if ((r = adjoint_pass(c, handles, nh, 1))) { hipStreamSynchronize(s); cleanup(); return r; }
    pm_launch_fix_combine(s, n3, c->d_adj_fix, c->d_adj_out, mw, w);
    // two classical Gram-Schmidt passes against v_0..v_j; alpha_j = v_j . w
    // collects the v_j coefficients of both passes
    for (int pass = 0; pass < 2; pass++) {
      double* cf = coef + (size_t)pass * (m + 1);
      pm_launch_dots(s, n3, j + 1, V, w, part, cf, j, alpha + j, pass == 0, 0);
      pm_launch_gs_update(s, n3, j + 1, V, cf, w);
    }
    // beta_j = |w|, v_{j+1} = w / beta_j
    pm_launch_dots(s, n3, 1, w, w, part, coef, 0, beta + j, 1, 1);
    pm_launch_normalise(s, n3, w, beta + j, 1e-300, V + (size_t)(j + 1) * n3);
  }
  std::vector<double> abh(2 * (size_t)m);
  e = hipMemcpyAsync(abh.data(), ab, abh.size() * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) { cleanup(); return fail(c, PM_ERR_DEVICE, "HIP error %s in the Lanczos steps", hipGetErrorString(e)); }
  if ((r = check_overflow(c))) { cleanup(); return r; }
  // tridiagonal T (truncated where an invariant subspace ended the recurrence)
  int mt = m;
  const double bscale = std::fabs(abh[0]) + 1e-300;
  for (int j = 0; j < m - 1; j++)
    if (abh[m + j] <= 1e-14 * bscale) { mt = j + 1; break; }
  std::vector<double> T((size_t)mt * mt, 0.0), d, z;
  for (int j = 0; j < mt; j++) {
    T[(size_t)j * mt + j] = abh[j];
    if (j + 1 < mt) T[(size_t)j * mt + j + 1] = T[(size_t)(j + 1) * mt + j] = abh[m + j];
  }
  jacobi_eigen(mt, T, d, z);
  std::vector<int> ord(mt);
  for (int i = 0; i < mt; i++) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  const double blast = abh[m + mt - 1];
  for (int i = 0; i < m; i++) {
    ritz_values[i] = i < mt ? d[ord[i]] : std::nan("");
    if (residuals) residuals[i] = i < mt ? std::fabs(blast * z[(size_t)(mt - 1) * mt + ord[i]]) : std::nan("");
  }
  if (alpha_beta) std::copy(abh.begin(), abh.end(), alpha_beta);
  if (nvec > 0) {
    const int ny = std::min(nvec, mt);
    std::vector<double> Sh((size_t)ny * mt);
    for (int i = 0; i < ny; i++)
      for (int j = 0; j < mt; j++) Sh[(size_t)i * mt + j] = z[(size_t)j * mt + ord[i]];
    e = hipMemcpyAsync(S, Sh.data(), Sh.size() * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) pm_launch_ritz(s, n3, mt, V, S, ny, Y);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(ritz_vectors, Y, (size_t)ny * bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (int i = ny; i < nvec; i++)
      std::fill(ritz_vectors + (size_t)i * n3, ritz_vectors + (size_t)(i + 1) * n3, 0.0);
    if (e != hipSuccess) { cleanup(); return fail(c, PM_ERR_DEVICE, "HIP error %s assembling the Ritz vectors", hipGetErrorString(e)); }
  }
  cleanup();
  return PM_OK;
}

int pm_set_option(pm_context* c, int option, int value) {
  if (!c) return PM_ERR_ARGUMENT;
  c->gen++;
  if (option == PM_OPT_CLUSTER_SIZE) {
    if (value != 0 && value != 4 && value != 8)
      return fail(c, PM_ERR_ARGUMENT, "cluster size must be 0 (automatic), 4 or 8 (got %d)", value);
    c->m_override = value;
    c->md_forces_valid = false;
    c->gen++;
    return PM_OK;
  }
  if (option == PM_OPT_GRAPH) {
    if (value != 0 && value != 1) return fail(c, PM_ERR_ARGUMENT, "graph option must be 0 or 1 (got %d)", value);
    c->graph_option = value;
    return PM_OK;
  }
  if (option == PM_OPT_LIST_SKIN) {
    if (value < -1 || value > 3000)
      return fail(c, PM_ERR_ARGUMENT, "list skin must be -1 (automatic), 0 (off) or 1..3000 (1e-3 A; got %d)", value);
    c->skin_milli = value;
    c->list_stale = true;
    c->md_forces_valid = false;
    return PM_OK;
  }
  if (option == PM_OPT_LIST_TEST) {
    if (value < -1 || value > 1)
      return fail(c, PM_ERR_ARGUMENT, "list test must be -1 (automatic), 0 (bounding box where it holds) or 1 (exact; got %d)", value);
    c->list_test_option = value;
    c->list_stale = true;
    c->md_forces_valid = false;
    return PM_OK;
  }
  if (option == PM_OPT_PRUNE_SKIN) {
    if (value < -1 || value > 3000)
      return fail(c, PM_ERR_ARGUMENT, "inner-list skin must be -1 (automatic), 0 (off) or 1..3000 (1e-3 A; got %d)", value);
    c->prune_milli = value;
    c->list_stale = true;
    c->md_forces_valid = false;
    return PM_OK;
  }
  if (option == PM_OPT_PME_PRIORITY) {
    if (value < -1 || value > 3)
      return fail(c, PM_ERR_ARGUMENT, "PME priority must be -1 (automatic) or 0..3 (got %d)", value);
    c->pme_prio_option = value;
    c->gen++;   // captured steps hold the launch arguments
    return PM_OK;
  }
  if (option == PM_OPT_PME_SPREAD) {
    if (value < -1 || value > 1)
      return fail(c, PM_ERR_ARGUMENT, "PME spread option must be -1 (automatic), 0 (cell bricks) or 1 (fused with the z transform; got %d)", value);
    c->pme_spread_option = value == 0 ? 0 : 1;
    c->gen++;   // captured steps hold the launch arguments
    return PM_OK;
  }
  if (option == PM_OPT_PAIR_LANES) {
    if (value < -1 || value > 1)
      return fail(c, PM_ERR_ARGUMENT, "pair lane option must be -1 (automatic), 0 (a row per entry) or 1 (a lane per entry; got %d)", value);
    c->pair_lanes_option = value == 0 ? 0 : 1;
    c->gen++;   // captured steps hold the launch arguments
    return PM_OK;
  }
  if (option == PM_OPT_PME_FFT) {
    if (value < 0 || value > 2)
      return fail(c, PM_ERR_ARGUMENT, "PME FFT option must be 0 (automatic), 1 (separate passes) or 2 (plane kernel; got %d)", value);
    c->pme_fft_option = value;
    c->md_forces_valid = false;
    return PM_OK;
  }
  return fail(c, PM_ERR_ARGUMENT, "unknown option %d", option);
}

int pm_get_option(pm_context* c, int option, int* value) {
  if (!c || !value) return PM_ERR_ARGUMENT;
  if (option == PM_OPT_CLUSTER_SIZE) {
    *value = c->have_topo && c->have_box ? choose_m(c) : c->m_override;
    return PM_OK;
  }
  if (option == PM_OPT_GRAPH) {
    *value = c->graph_option;
    return PM_OK;
  }
  if (option == PM_OPT_LIST_SKIN) {
    *value = (int)std::lround(list_skin(c) * 1000.0);
    return PM_OK;
  }
  if (option == PM_OPT_PME_PRIORITY) {
    *value = pme_priority(c);
    return PM_OK;
  }
  if (option == PM_OPT_PRUNE_SKIN) {
    *value = (int)std::lround(prune_skin(c) * 1000.0);
    return PM_OK;
  }
  if (option == PM_OPT_LIST_TEST) {
    *value = list_exact(c) ? 1 : 0;
    return PM_OK;
  }
  if (option == PM_OPT_PME_SPREAD) {
    *value = c->pme_spread_option;
    return PM_OK;
  }
  if (option == PM_OPT_PAIR_LANES) {
    *value = pair_lanes(c);
    return PM_OK;
  }
  if (option == PM_OPT_PME_FFT) {
    *value = c->pme_fft_option == 2 ? 2 : 1;
    if (c->pme_nx > 0 && *value == 2 && !pm_pme_plane_fits(c->pme_nx, c->pme_ny, c->precision, c->lds_max))
      *value = 1;
    return PM_OK;
  }
  return fail(c, PM_ERR_ARGUMENT, "unknown option %d", option);
}

int pm_synchronize(pm_context* c) {
  if (!c) return PM_ERR_ARGUMENT;
  int r;
  if ((r = set_device(c))) return r;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream2));
  return check_overflow(c);
}

}  // extern "C"

// ------------------------------------------------------------------ batches
// Many independent systems (replicas), each in its own context, advanced
// together: their MD steps are captured into one HIP graph (every replica on
// its own streams, joined on the batch's root stream) and replayed; replica
// exchange swaps bath temperatures between neighbours of a temperature ladder.
// One MD step of every replica as one launch per kernel: the recorded kernel
// sequence, shared by all replicas, with device tables of their arguments.
struct BatchPlan {
  struct Item {
    int kind, variant;
    unsigned threads, lds, total, argbytes;
    size_t args_off, pre_off;
  };
  std::vector<Item> items;
  unsigned char* d_buf = nullptr;
  std::vector<unsigned long long> gen;
  hipGraphExec_t exec = nullptr;
  int steps = 0;
  bool unsupported = false;   // the replicas at generations `gen` have no batched form
};

struct pm_batch {
  std::vector<pm_context*> ctx;
  hipStream_t root = nullptr;
  hipEvent_t ev_root = nullptr;
  std::vector<hipEvent_t> ev_done;
  hipGraphExec_t exec = nullptr;
  int exec_steps = 0;
  std::vector<unsigned long long> exec_gen;
  std::vector<int> ladder;    // ladder[k]: replica holding temperature k of the ladder
  std::vector<double> temps;  // the ladder temperatures, as given at creation
  std::string err;
  int mode = PM_BATCH_AUTO;   // PM_BATCH_AUTO: batched launches when the replicas allow them
  BatchPlan plan;             // steps without energies (captured, G steps per replay)
  BatchPlan eplan;            // the step with energies
  bool batched_last = false;  // the last pm_batch_run used batched launches
  std::string why_chains;     // why the replicas ran as chains
  int exchange_velocities = PM_EXCHANGE_RESCALE;   // what an accepted swap does to the velocities
};

namespace {

int bfail(pm_batch* b, int code, const char* msg, const pm_context* c = nullptr) {
  b->err = msg;
  if (c && !c->err.empty()) b->err += std::string(": ") + c->err;
  return code;
}

void batch_drop_graph(pm_batch* b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  b->exec_steps = 0;
}

void plan_free(BatchPlan& P);
void batch_drop_plan(pm_batch* b) {
  plan_free(b->plan);
  plan_free(b->eplan);
}

// the root stream after everything queued on the replicas' streams
int batch_join_replicas(pm_batch* b) {
  for (size_t k = 0; k < b->ctx.size(); k++) {
    pm_context* c = b->ctx[k];
    if (hipEventRecord(b->ev_done[k], c->stream) != hipSuccess || hipStreamWaitEvent(b->root, b->ev_done[k], 0) != hipSuccess)
      return bfail(b, PM_ERR_DEVICE, "stream join failed");
  }
  return PM_OK;
}

// launch `steps` steps of a plan's kernels directly on the root stream (or
// into the capture under way)
bool batch_plan_launch(pm_batch* b, const BatchPlan& P, int steps) {
  const int R = (int)b->ctx.size();
  for (int u = 0; u < steps; u++)
    for (const BatchPlan::Item& it : P.items) {
      if (it.total == 0) continue;
      const void* t = P.d_buf + it.args_off;
      const int* pre = reinterpret_cast<const int*>(P.d_buf + it.pre_off);
      if (!(pm_launch_batched_md(b->root, it.kind, it.variant, t, pre, R, it.total, it.threads, it.lds) ||
            pm_launch_batched_nb(b->root, it.kind, it.variant, t, pre, R, it.total, it.threads, it.lds) ||
            pm_launch_batched_bonded(b->root, it.kind, it.variant, t, pre, R, it.total, it.threads, it.lds)))
        return false;
    }
  return true;
}

// Record one step (with or without energies) of every replica into plan P:
// PM_OK, or PM_ERR_UNSUPPORTED (why_chains says why) when the replicas'
// kernel sequences differ or hold a kernel without a batched form.  With
// graph_steps > 0 the plan also holds a captured graph of that many steps.
int batch_make_plan(pm_batch* b, bool energy, int graph_steps, BatchPlan& P) {
  const int R = (int)b->ctx.size();
  std::vector<std::vector<PmRecord>> rec(R);
  for (int k = 0; k < R; k++) {
    pm_context* c = b->ctx[k];
    for (int h : c->md_handles)
      if (c->forces[h].pme) { b->why_chains = "a replica has a PME force"; return PM_ERR_UNSUPPORTED; }
    pm_set_recorder(&rec[k]);
    const int r = md_unit(c, 1, energy);
    pm_set_recorder(nullptr);
    c->md_counter -= 2;   // nothing ran: the plan's launches advance the device counter
    if (r) return bfail(b, r, "replica step failed", c);
  }
  const size_t L = rec[0].size();
  for (int k = 0; k < R; k++) {
    if (rec[k].size() != L) { b->why_chains = "the replicas' kernel sequences differ in length"; return PM_ERR_UNSUPPORTED; }
    for (size_t i = 0; i < L; i++) {
      const PmRecord &x = rec[0][i], &y = rec[k][i];
      if (y.kind <= 0) { b->why_chains = "a replica step holds a kernel without a batched form"; return PM_ERR_UNSUPPORTED; }
      if (x.kind != y.kind || x.variant != y.variant || x.threads != y.threads || x.argbytes != y.argbytes) {
        b->why_chains = "the replicas' kernel sequences differ (forces, precision or cluster size)";
        return PM_ERR_UNSUPPORTED;
      }
      if (y.kind == PM_REC_PAIR && ((y.variant >> 8) & 1)) {
        b->why_chains = "pair kernels of 4-atom clusters have no batched form";
        return PM_ERR_UNSUPPORTED;
      }
    }
  }
  // argument tables and block prefixes, 256-byte aligned
  std::vector<unsigned char> host;
  auto align = [&]() { host.resize((host.size() + 255) / 256 * 256, 0); };
  for (size_t i = 0; i < L; i++) {
    BatchPlan::Item it{};
    it.kind = rec[0][i].kind;
    it.variant = rec[0][i].variant;
    it.threads = rec[0][i].threads;
    it.argbytes = rec[0][i].argbytes;
    align();
    it.args_off = host.size();
    for (int k = 0; k < R; k++) host.insert(host.end(), rec[k][i].args, rec[k][i].args + it.argbytes);
    align();
    it.pre_off = host.size();
    std::vector<int> pre(R + 1, 0);
    unsigned lds = 0;
    for (int k = 0; k < R; k++) {
      pre[k + 1] = pre[k] + (int)rec[k][i].blocks;
      lds = std::max(lds, rec[k][i].lds);
    }
    it.total = (unsigned)pre[R];
    it.lds = lds;
    const unsigned char* pb = reinterpret_cast<const unsigned char*>(pre.data());
    host.insert(host.end(), pb, pb + pre.size() * sizeof(int));
    P.items.push_back(it);
  }
  if (hipMalloc((void**)&P.d_buf, std::max<size_t>(host.size(), 256)) != hipSuccess ||
      hipMemcpy(P.d_buf, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess)
    return bfail(b, PM_ERR_DEVICE, "batched plan upload failed");
  if (graph_steps > 0) {
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(b->root, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return bfail(b, PM_ERR_DEVICE, "hipStreamBeginCapture failed");
    const bool ok = batch_plan_launch(b, P, graph_steps);
    hipError_t e = hipStreamEndCapture(b->root, &g);
    if (!ok) {
      if (g) hipGraphDestroy(g);
      b->why_chains = "a recorded kernel has no batched launcher";
      return PM_ERR_UNSUPPORTED;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&P.exec, g, nullptr, nullptr, 0);
    if (g) hipGraphDestroy(g);
    if (e != hipSuccess) { P.exec = nullptr; return bfail(b, PM_ERR_DEVICE, "batched graph failed"); }
    P.steps = graph_steps;
  }
  for (pm_context* c : b->ctx) P.gen.push_back(c->gen);
  return PM_OK;
}

void plan_free(BatchPlan& P) {
  if (P.exec) hipGraphExecDestroy(P.exec);
  if (P.d_buf) hipFree(P.d_buf);
  P = BatchPlan();
}

// both plans of a batch (steps without energies replayed G at a time, the
// energy step); PM_ERR_UNSUPPORTED leaves the batch on per-replica chains
int batch_build_plan(pm_batch* b, int G) {
  batch_drop_plan(b);
  int r = batch_make_plan(b, false, G, b->plan);
  if (r == PM_OK) r = batch_make_plan(b, true, 0, b->eplan);
  if (r != PM_OK) {
    const std::string why = b->why_chains;
    batch_drop_plan(b);
    b->why_chains = why;
  }
  return r;
}

// Langevin steps without energies for every replica: each replica's steps
// form one chain on its own stream, forked from and joined to the batch's
// root stream once per call
int batch_units(pm_batch* b, int steps) {
  if (hipEventRecord(b->ev_root, b->root) != hipSuccess) return bfail(b, PM_ERR_DEVICE, "event record failed");
  for (size_t k = 0; k < b->ctx.size(); k++) {
    pm_context* c = b->ctx[k];
    if (hipStreamWaitEvent(c->stream, b->ev_root, 0) != hipSuccess)
      return bfail(b, PM_ERR_DEVICE, "stream wait failed");
    for (int u = 0; u < steps; u++) {
      int r = md_unit(c, 1, false);
      if (r) return bfail(b, r, "replica step failed", c);
    }
    if (hipEventRecord(b->ev_done[k], c->stream) != hipSuccess) return bfail(b, PM_ERR_DEVICE, "event record failed");
    if (hipStreamWaitEvent(b->root, b->ev_done[k], 0) != hipSuccess)
      return bfail(b, PM_ERR_DEVICE, "stream wait failed");
  }
  return PM_OK;
}

// uniform deviate in [0, 1) of exchange attempt `attempt`, ladder pair `k`
double rem_uniform(unsigned long long seed, unsigned long long attempt, unsigned long long k) {
  auto mix = [](unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  unsigned long long h = mix(seed ^ mix(attempt * 0x100000001B3ull + 2 * k + 1));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace

extern "C" {

int pm_batch_create(pm_context* const* contexts, int nrep, pm_batch** out) {
  if (!contexts || nrep <= 0 || !out) return PM_ERR_ARGUMENT;
  *out = nullptr;
  for (int k = 0; k < nrep; k++) {
    pm_context* c = contexts[k];
    if (!c) return PM_ERR_ARGUMENT;
    if (c->md_handles.empty() || c->md_integrator != PM_INTEGRATOR_LANGEVIN) return PM_ERR_STATE;
    if (c->device != contexts[0]->device) return PM_ERR_ARGUMENT;
  }
  pm_batch* b = new pm_batch();
  b->ctx.assign(contexts, contexts + nrep);
  if (hipSetDevice(contexts[0]->device) != hipSuccess || hipStreamCreate(&b->root) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_root, hipEventDisableTiming) != hipSuccess) {
    delete b;
    return PM_ERR_DEVICE;
  }
  b->ev_done.resize(nrep, nullptr);
  for (int k = 0; k < nrep; k++) hipEventCreateWithFlags(&b->ev_done[k], hipEventDisableTiming);
  for (int k = 0; k < nrep; k++) {
    b->ladder.push_back(k);
    b->temps.push_back(contexts[k]->md_temp);
  }
  *out = b;
  return PM_OK;
}

void pm_batch_destroy(pm_batch* b) {
  if (!b) return;
  hipSetDevice(b->ctx[0]->device);
  hipStreamSynchronize(b->root);
  for (pm_context* c : b->ctx) hipStreamSynchronize(c->stream);
  batch_drop_graph(b);
  batch_drop_plan(b);
  for (hipEvent_t e : b->ev_done) hipEventDestroy(e);
  hipEventDestroy(b->ev_root);
  hipStreamDestroy(b->root);
  delete b;
}

const char* pm_batch_last_error(const pm_batch* b) { return b ? b->err.c_str() : "null batch"; }

int pm_batch_run(pm_batch* b, int nsteps) {
  if (!b) return PM_ERR_ARGUMENT;
  if (nsteps < 1) return PM_OK;
  if (hipSetDevice(b->ctx[0]->device) != hipSuccess) return bfail(b, PM_ERR_DEVICE, "hipSetDevice failed");
  int r;
  for (pm_context* c : b->ctx) {
    if (c->timing_mask) return bfail(b, PM_ERR_STATE, "kernel timing is on in a replica context");
    if (!c->md_forces_valid) {
      if ((r = evaluate_group(c, c->md_handles, 0, false))) return bfail(b, r, "initial forces failed", c);
      c->md_forces_valid = true;
    }
  }
  if ((r = batch_join_replicas(b))) return r;
  // every step but the last from a captured graph of up to 8 steps
  const int count = nsteps - 1;
  int done = 0;
  // graphs always hold 8 steps (a shorter run captures one for the next
  // runs and launches its steps directly), so runs of different lengths
  // never capture again
  const int G = 8;
  b->batched_last = false;
  if (count >= 2 && b->mode == PM_BATCH_AUTO) {
    // one launch per kernel over all replicas, replayed in steps of G
    bool same_gen = b->plan.gen.size() == b->ctx.size();
    for (size_t k = 0; same_gen && k < b->ctx.size(); k++) same_gen = b->plan.gen[k] == b->ctx[k]->gen;
    const bool stale = !same_gen || (!b->plan.unsupported && (!b->plan.exec || b->plan.steps != G));
    bool usable = !stale && !b->plan.unsupported;
    if (stale) {
      if ((r = batch_units(b, 1))) return r;   // every buffer exists and every layout is current
      done = 1;
      r = batch_build_plan(b, G);
      if (r == PM_OK) usable = true;
      else if (r != PM_ERR_UNSUPPORTED) return r;
      else {
        b->plan.unsupported = true;
        b->plan.gen.clear();
        for (pm_context* c : b->ctx) b->plan.gen.push_back(c->gen);
      }
    }
    if (usable) {
      b->batched_last = true;
      while (count - done >= G) {
        if (hipGraphLaunch(b->plan.exec, b->root) != hipSuccess) return bfail(b, PM_ERR_DEVICE, "hipGraphLaunch failed");
        for (pm_context* c : b->ctx) c->md_counter += 2ull * G;
        done += G;
      }
      // the remaining steps and the step with energies as direct batched launches
      if (!batch_plan_launch(b, b->plan, count - done) || !batch_plan_launch(b, b->eplan, 1))
        return bfail(b, PM_ERR_DEVICE, "batched launch failed");
      for (pm_context* c : b->ctx) c->md_counter += 2ull * (count - done + 1);
      if (hipGetLastError() != hipSuccess) return bfail(b, PM_ERR_DEVICE, "batched launch failed");
      // the replicas' streams after the batch (later per-replica work)
      if (hipEventRecord(b->ev_root, b->root) != hipSuccess) return bfail(b, PM_ERR_DEVICE, "event record failed");
      for (pm_context* c : b->ctx)
        if (hipStreamWaitEvent(c->stream, b->ev_root, 0) != hipSuccess) return bfail(b, PM_ERR_DEVICE, "stream wait failed");
      return PM_OK;
    }
  }
  if (count - done >= 2 && !b->batched_last) {
    bool stale = !b->exec || b->exec_steps != G || b->exec_gen.size() != b->ctx.size();
    for (size_t k = 0; !stale && k < b->ctx.size(); k++) stale = b->exec_gen[k] != b->ctx[k]->gen;
    if (stale) {
      batch_drop_graph(b);
      if ((r = batch_units(b, 1))) return r;   // every buffer exists before the capture
      done += 1;
      std::vector<unsigned long long> gen0;
      for (pm_context* c : b->ctx) gen0.push_back(c->gen);
      hipGraph_t g = nullptr;
      if (hipStreamBeginCapture(b->root, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return bfail(b, PM_ERR_DEVICE, "hipStreamBeginCapture failed");
      int rc = batch_units(b, G);
      hipError_t e = hipStreamEndCapture(b->root, &g);
      // the capture queued nothing: its steps run when the graph is launched
      for (pm_context* c : b->ctx) c->md_counter -= 2ull * G;
      if (rc) { if (g) hipGraphDestroy(g); return rc; }
      if (e != hipSuccess) return bfail(b, PM_ERR_DEVICE, "hipStreamEndCapture failed");
      e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) { b->exec = nullptr; return bfail(b, PM_ERR_DEVICE, "hipGraphInstantiate failed"); }
      for (size_t k = 0; k < b->ctx.size(); k++)
        if (b->ctx[k]->gen != gen0[k]) { batch_drop_graph(b); return bfail(b, PM_ERR_STATE, "internal: buffers changed during capture"); }
      b->exec_steps = G;
      b->exec_gen = gen0;
    }
    while (count - done >= G) {
      if (hipGraphLaunch(b->exec, b->root) != hipSuccess) return bfail(b, PM_ERR_DEVICE, "hipGraphLaunch failed");
      for (pm_context* c : b->ctx) c->md_counter += 2ull * G;
      done += G;
    }
  }
  if (done < count && (r = batch_units(b, count - done))) return r;
  // the last step with energies
  if (hipEventRecord(b->ev_root, b->root) != hipSuccess) return bfail(b, PM_ERR_DEVICE, "event record failed");
  for (pm_context* c : b->ctx) {
    if (hipStreamWaitEvent(c->stream, b->ev_root, 0) != hipSuccess) return bfail(b, PM_ERR_DEVICE, "stream wait failed");
    if ((r = md_unit(c, 1, true))) return bfail(b, r, "replica step failed", c);
  }
  // list capacity, non-finite positions and force range of every replica
  for (pm_context* c : b->ctx)
    if ((r = check_overflow(c))) return bfail(b, r, "replica step failed", c);
  return PM_OK;
}

int pm_batch_exchange(pm_batch* b, int attempt, unsigned long long seed, int* accepted, double* deltas) {
  if (!b || attempt < 0) return PM_ERR_ARGUMENT;
  const int R = (int)b->ctx.size();
  if (hipSetDevice(b->ctx[0]->device) != hipSuccess) return bfail(b, PM_ERR_DEVICE, "hipSetDevice failed");
  // potential energy of every replica at its last step (ScalarStructure::potentialEnergy)
  std::vector<double> epot(R, 0.0);
  for (int k = 0; k < R; k++) {
    pm_context* c = b->ctx[k];
    double e[PM_NE];
    if (hipMemcpyAsync(e, c->d_energy, sizeof(e), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      return bfail(b, PM_ERR_DEVICE, "energy copy failed");
    for (int u = 0; u < PM_E_VIR; u++) epot[k] += e[u];
  }
  // neighbour pairs (k, k+1) of the ladder with k of the attempt's parity
  // (getNextAddress, applications/rem-app/rem.cpp:949-957); the Metropolis
  // test on calcDelta (:941-946, canDoEvenSwitch :841-880) swaps the two
  // replicas' bath temperatures; with PM_EXCHANGE_RESCALE their velocities
  // follow by sqrt(T_new / T_old) (doSwitch's rescaling branch, :930-935),
  // with PM_EXCHANGE_KEEP they are left for the caller to redraw (doSwitch
  // as the reference runs it, redrawVelocities = 1, :912, :923-928)
  const double kb = 0.001987191;   // Constant::BOLTZMANN (base/pmconstants.cpp:74)
  for (int k = attempt & 1; k + 1 < R; k += 2) {
    const int ra = b->ladder[k], rb = b->ladder[k + 1];
    const double ta = b->temps[k], tb = b->temps[k + 1];
    const double delta = (1 / kb / ta - 1 / kb / tb) * (epot[rb] - epot[ra]);
    const double w = delta <= 0 ? 1.0 : std::exp(-delta);
    const bool acc = rem_uniform(seed, (unsigned long long)attempt, (unsigned long long)k) <= w;
    if (deltas) deltas[k] = delta;
    if (accepted) accepted[k] = acc ? 1 : 0;
    if (!acc) continue;
    std::swap(b->ladder[k], b->ladder[k + 1]);
    const int repl[2] = {ra, rb};
    const double tnew[2] = {tb, ta}, told[2] = {ta, tb};
    for (int u = 0; u < 2; u++) {
      pm_context* c = b->ctx[repl[u]];
      c->md_temp = tnew[u];
      if (hipMemcpyAsync(&c->d_md->temperature, &c->md_temp, sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return bfail(b, PM_ERR_DEVICE, "temperature copy failed");
      if (b->exchange_velocities == PM_EXCHANGE_RESCALE)
        pm_launch_scale(c->stream, 3 * c->n, std::sqrt(tnew[u] / told[u]), c->d_vel);
      if (hipStreamSynchronize(c->stream) != hipSuccess) return bfail(b, PM_ERR_DEVICE, "exchange failed");
    }
  }
  return PM_OK;
}

int pm_batch_set_exchange_velocities(pm_batch* b, int mode) {
  if (!b || (mode != PM_EXCHANGE_RESCALE && mode != PM_EXCHANGE_KEEP)) return PM_ERR_ARGUMENT;
  b->exchange_velocities = mode;
  return PM_OK;
}

int pm_batch_set_mode(pm_batch* b, int mode) {
  if (!b || (mode != PM_BATCH_AUTO && mode != PM_BATCH_CHAINS)) return PM_ERR_ARGUMENT;
  b->mode = mode;
  return PM_OK;
}

int pm_batch_info(pm_batch* b, int* batched, char* why, int why_len) {
  if (!b) return PM_ERR_ARGUMENT;
  if (batched) *batched = b->batched_last ? 1 : 0;
  if (why && why_len > 0) {
    std::snprintf(why, (size_t)why_len, "%s", b->batched_last ? "" : b->why_chains.c_str());
  }
  return PM_OK;
}

int pm_batch_ladder(pm_batch* b, int* ladder, double* temps) {
  if (!b) return PM_ERR_ARGUMENT;
  for (size_t k = 0; k < b->ctx.size(); k++) {
    if (ladder) ladder[k] = b->ladder[k];
    if (temps) temps[k] = b->temps[k];
  }
  return PM_OK;
}

}  // extern "C"
