// engine.hip — the per-step force path on one MI355X: buffers, launch order,
// step graphs and the C entry points of include/mdrt_hip.h (group 2).
//
// Step order (the reference's, 08-gpu.md:112-121):
//   integrate (second half of a step and first half of the next: half
//   kicks, drift, SETTLE, f32 positions by place, f64 positions by atom for
//   PME, the validity test), then two branches joined before the next one:
//   - the list branch: nb_bin / nb_sort / nb_search (no-ops unless the
//     validity test failed), nb_pairs (LJ + erfc Ewald, 64-bit force words;
//     its first plane takes the excluded-pair correction and 1-4 pairs);
//   - the PME branch, in atom order: pme_bin, pme_spread_rows (spread + z
//     transform), fft_plane_conv (y and x transforms, influence, inverse),
//     pme_gather_rows (inverse z transform + gather into its own words).
// The branches run on two streams with default priorities, forked and
// joined by events, so a run of steps captures into one HIP graph and
// replays without host decisions. Launch shapes (segment length, PME
// bricks, plane transform) are runtime choices the launch-shape trial
// (autotune) times on the system.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <unordered_map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mdrt_hip.h"
#include "integrate.hip"
#include "nbnxm.hip"
#include "bonded.hip"
#include "pairs.hip"
#include "pme.hip"
#include "pme_mfma.hip"

namespace mdhip {

static thread_local std::string g_last_error;

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHK(expr)                                                               \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      throw ::mdhip::HipError(std::string(#expr) + ": " + hipGetErrorString(_e));  \
  } while (0)

// Setup allocations are cleared and uploaded through the null stream and
// then waited for on the whole device: the engine's own stream is
// non-blocking and would not otherwise be ordered after them.
template <typename T>
static T *dalloc(size_t count) {
  void *p = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc(&p, count * sizeof(T)));
  HIPCHK(hipMemset(p, 0, count * sizeof(T)));
  HIPCHK(hipDeviceSynchronize());
  return static_cast<T *>(p);
}

template <typename T>
static void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

template <typename T>
static T *dupload(const std::vector<T> &v) {
  T *p = dalloc<T>(v.size());
  if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  HIPCHK(hipDeviceSynchronize());
  return p;
}

static std::vector<int> radices(int K) {
  std::vector<int> r;
  int rest = K;
  for (int f : {8, 4, 2, 3, 5, 7})
    while (rest % f == 0) {
      r.push_back(f);
      rest /= f;
    }
  if (rest != 1 || (int)r.size() > kMaxRadices)
    throw HipError("FFT length " + std::to_string(K) + " has a factor other than 2, 3, 5, 7");
  return r;
}

// Twiddles exp(-2 pi i m / K), m < K, rounded once from f64.
template <typename T>
struct FftTables {
  int K = 0;
  cpx<T> *tw = nullptr;
  FftAxis axis{};
  void make(int k) {
    K = k;
    std::vector<cpx<T>> h(K);
    for (int m = 0; m < K; ++m) {
      const double a = -2.0 * M_PI * (double)m / (double)K;
      h[m] = {(T)cos(a), (T)sin(a)};
    }
    tw = dupload(h);
    std::vector<int> r = radices(K);
    axis.K = K;
    axis.nrad = (int)r.size();
    for (int i = 0; i < axis.nrad; ++i) axis.rad[i] = r[i];
    axis.tw = tw;
  }
  void release() { dfree(tw); }
};

// The 3D real transform of a grid: passes z (rows), y (columns), x.
template <typename T>
struct Fft3D {
  int K[3] = {0, 0, 0};
  FftTables<T> ax[3];
  void make(const int *k) {
    for (int a = 0; a < 3; ++a) {
      K[a] = k[a];
      ax[a].make(k[a]);
    }
  }
  void release() {
    for (auto &a : ax) a.release();
  }
  static constexpr int TL = tile_lines<T>();
  size_t lds(int k) const { return (size_t)2 * TL * k * sizeof(cpx<T>); }
  int H() const { return K[2] / 2 + 1; }
  // real (or 64-bit words) -> half spectrum
  void forward(const void *src, bool fixed, cpx<T> *cplx, hipStream_t s) {
    const int nrows = K[0] * K[1];
    const int blocks = (nrows + 2 * TL - 1) / (2 * TL);
    if constexpr (sizeof(T) == 4) {
      if (fixed)
        fft_rows_r2c<T, true><<<blocks, 256, lds(K[2]), s>>>(src, cplx, nrows, ax[2].axis);
      else
        fft_rows_r2c<T, false><<<blocks, 256, lds(K[2]), s>>>(src, cplx, nrows, ax[2].axis);
    } else {
      fft_rows_r2c<T, false><<<blocks, 256, lds(K[2]), s>>>(src, cplx, nrows, ax[2].axis);
    }
    const int tiles = (H() + TL - 1) / TL;
    fft_cols_y<T><<<K[0] * tiles, 256, lds(K[1]), s>>>(cplx, K[1], H(), ax[1].axis, -1);
  }
  void x_pass(cpx<T> *cplx, int mode, const ConvTables &ct, double *epart, hipStream_t s) {
    const int tiles = (H() + TL - 1) / TL;
    fft_x_conv<T><<<K[1] * tiles, 256, lds(K[0]), s>>>(cplx, K[0], K[1], H(), ax[0].axis,
                                                         ct, mode, epart);
  }
  int x_blocks() const { return K[1] * ((H() + TL - 1) / TL); }
  void backward(cpx<T> *cplx, T *dst, hipStream_t s) {
    const int tiles = (H() + TL - 1) / TL;
    fft_cols_y<T><<<K[0] * tiles, 256, lds(K[1]), s>>>(cplx, K[1], H(), ax[1].axis, +1);
    const int nrows = K[0] * K[1];
    const int blocks = (nrows + 2 * TL - 1) / (2 * TL);
    fft_rows_c2r<T><<<blocks, 256, lds(K[2]), s>>>(cplx, dst, nrows, ax[2].axis, 0);
  }
};

static inline int blocks_for(size_t n, int b = 256) { return (int)((n + b - 1) / b); }

// The potential energy: every term but the kinetic energy.
inline double potential_of(const double *E) {
  double u = 0.0;
  for (int k = 0; k < MDRT_NUM_ENERGIES; ++k)
    if (k != MDRT_E_KINETIC) u += E[k];
  return u;
}

struct Engine {
  // ---- topology ----
  int n = 0, ntypes = 0, nunits = 0;
  double box[3] = {0, 0, 0};
  double cutoff = 0, switch_from = 0, skin = 0.1, beta = 0;
  double prune_skin = 0;       // > 0: the pair loop runs over a pruned list
  int lj_modifier = 0, coulomb_shift = 0, dispersion = 0;
  int pme_grid[3] = {0, 0, 0}, pme_order = 4;
  std::vector<double> h_q, h_mass, h_sigma, h_eps;
  std::vector<int> h_type;
  std::vector<int> h_excl_pairs;  // flat i<j
  double e_self = 0, e_disp = 0;

  hipStream_t stream = nullptr;
  // The PME chain runs on a second stream, concurrently with the pair loop
  // (both add integer force words, so the order does not matter); events
  // fork it after the list build and join it before the integration.
  hipStream_t side = nullptr;
  hipStream_t pstream = nullptr;   // where the PME launches go
  hipStream_t lstream = nullptr;   // where the list build and the pair loop go
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;

  // ---- state (atom order) ----
  double *d_x = nullptr, *d_v = nullptr, *d_xref = nullptr, *d_shift = nullptr;
  double *d_q = nullptr, *d_inv_mass = nullptr;
  int *d_type = nullptr, *d_place_of = nullptr;
  double *d_save_x = nullptr, *d_save_v = nullptr;

  // ---- places, clusters, list ----
  GridParams g{};
  int max_places = 0;
  int *d_atom_at = nullptr, *d_type_at = nullptr;
  float4 *d_xq = nullptr;
  float4 *d_gorig = nullptr;  // origin of each group of 32 places (xq is relative to it)
  // Group origins for boxes with a half edge from kOriginHalfEdge up: there
  // the f32 spacing about the box centre (4.8e-7 nm from 4 nm) outweighs the
  // rounding of each pair's origin offset (about 1.2e-7 nm).  Frame batches
  // keep positions about the box centre (the adjoint pair loop reads them).
  static constexpr double kOriginHalfEdge = 4.0;
  bool origins_allowed = true;
  bool use_origins() const {
    return origins_allowed && 0.5 * std::max(box[0], std::max(box[1], box[2])) >= kOriginHalfEdge;
  }
  double4 *d_xpq = nullptr;   // f64 positions and charges in place order (PME)
  long long *d_F = nullptr;
  // PME in atom order (pme_atoms): the reciprocal part reads positions and
  // charges the integrator writes by atom and adds its force words to d_Fr
  // by atom, so it depends on no list build and runs beside it.  Frame
  // batches keep PME in place order.
  bool pme_atoms = true;
  double4 *d_xpa = nullptr;   // f64 positions and charges in atom order
  long long *d_Fr = nullptr;  // reciprocal force words in atom order
  int *d_pme_n = nullptr;     // [0] = n, the element count of pme_bin
  int *d_col_count = nullptr, *d_bucket = nullptr;
  float4 *d_bb_lo = nullptr, *d_bb_hi = nullptr;
  int *d_cluster_col = nullptr;
  uint32_t *d_real_mask = nullptr;
  int *d_col_cl_start = nullptr, *d_col_cl_count = nullptr;
  int *d_counts = nullptr;
  PairEntry *d_rows = nullptr;     // group rows, row_cap entries each
  int *d_row_count = nullptr;
  uint64_t *d_ex = nullptr;        // explicit masks, kGroup per slot
  PairEntry *d_rows_in = nullptr;  // pruned rows (same capacity)
  int *d_row_count_in = nullptr;
  double *d_xprune = nullptr;      // positions at the last prune
  int ex_cap = 0;
  int max_groups = 0, nseg = 0;    // group rows; pair-loop segments per row
  // entries per pair-loop segment (16, 24 or 48): short segments give one
  // system more waves to hide latency with; a batch of frames fills the
  // GPU with long ones (a launch shape, see autotune)
  int seg_len = 16, seg_len_cfg = 0;
  int *d_excl_start = nullptr, *d_excl_atoms = nullptr;
  int *d_flags = nullptr;
  double list_reach = 0;  // rlist of the current list configuration

  // ---- nonbonded ----
  NbParams nbp{};
  float2 *d_lj = nullptr;
  double *d_epair = nullptr;

  // ---- bonded tuples (sorted by kind) ----
  int ntuples = 0;
  int4 *d_tuple_atoms = nullptr;
  int *d_tuple_kind = nullptr;
  double4 *d_tuple_param = nullptr;
  double *d_etup = nullptr;                // kTupleKinds per tuple block
  static constexpr int kTupleBlocksMax = 64;
  int tuple_blocks() const {
    return ntuples && (terms & 8) ? std::min(blocks_for(ntuples), kTupleBlocksMax) : 0;
  }

  // ---- constraint clusters ----
  int nclusters = 0;
  double *d_clen = nullptr;                // each member's length to its centre
  int *d_cl_first = nullptr, *d_cl_code = nullptr;  // cluster units

  // ---- special pairs ----
  int nspecial = 0;
  int2 *d_special = nullptr;
  int *d_is14 = nullptr, *d_isexcl = nullptr;
  double *d_p14_sigma = nullptr, *d_p14_eps = nullptr, *d_p14_qq = nullptr;
  SpecialParams spp{};
  double *d_espec = nullptr;

  // ---- PME ----
  bool use_pme = false;
  PmeParams pmp{};
  Fft3D<float> fft;
  cpx<float> *d_cplx = nullptr;
  float *d_tab = nullptr;
  ConvTables ct{};
  double *d_erecip = nullptr;
  double *d_vrecip = nullptr;  // reciprocal virial per plane block, when asked (frames)
  // The y and x transforms of the planes: 0 = LDS Stockham FFT
  // (fft_plane_conv), 1 = DFT matrix products on MFMA in split BF16x3
  // (fft_plane_conv_mfma; plane edges 32 or 64).
  int pme_transform = 0;
  __bf16 *d_dft1 = nullptr, *d_dft2 = nullptr;

  // ---- integration ----
  int *d_unit_first = nullptr, *d_unit_size = nullptr;
  IntParams ip{};

  // ---- Langevin thermostat ----
  // friction > 0: the steps are g-BAOAB Langevin steps at `temperature`
  // (K) with the noise of (seed, step number, atom); 0: velocity Verlet.
  double lv_friction = 0.0, lv_temperature = 0.0;
  unsigned long long lv_seed = 0;
  long long *d_lstep = nullptr;   // [0] step number of the next step, [1] blocks done
  int64_t step_count = 0;         // steps taken: the step number of the next step
  void set_langevin(double temperature, double friction, unsigned long long seed) {
    if (friction < 0.0 || !(temperature >= 0.0)) throw HipError("Langevin: friction and temperature must not be negative");
    lv_friction = friction;
    lv_temperature = temperature;
    lv_seed = seed;
    if (friction > 0.0 && !d_lstep) d_lstep = dalloc<long long>(2);
    invalidate_graph();  // captured graphs hold the integration parameters
  }
  // ---- Monte Carlo barostat ----
  // Isotropic volume moves: the units' centres of mass scale with the box
  // (scale_units), the energy is evaluated at the new volume, and the move
  // is kept with the Metropolis probability of
  //   W = U1 - U0 + P (V1 - V0) - N_units kT ln(V1 / V0)
  // (the isothermal-isobaric ensemble of molecular centres; Eppenga and
  // Frenkel, Mol. Phys. 52, 1303 (1984); Chow and Ferguson, Comput. Phys.
  // Commun. 91, 283 (1995)).  The volume change is uniform in
  // [-dv, dv]; dv starts at 1 % of the volume and is adapted every 10
  // moves towards an acceptance between 25 % and 75 %.  The random numbers
  // are Philox words of (seed, step number, 0, stream 1 << 24).
  double mc_dv = 0.0;
  int64_t mc_tried = 0, mc_kept = 0, mc_window_tried = 0, mc_window_kept = 0;

  void set_box(const double *nb) {
    const double rl = cutoff + skin;
    for (int k = 0; k < 3; ++k)
      if (!(nb[k] > 2.0 * rl)) throw HipError("every box edge must exceed 2 (cutoff + skin)");
    for (int k = 0; k < 3; ++k) {
      box[k] = nb[k];
      nbp.box[k] = (float)box[k];
      nbp.box_lo[k] = (float)(box[k] - (double)nbp.box[k]);
      nbp.half_box[k] = (float)(0.5 * box[k]);
      spp.box[k] = box[k];
      ip.half_box[k] = 0.5 * box[k];
      pmp.box[k] = box[k];
      pmp.inv_box[k] = 1.0 / box[k];
    }
    host_constants();
    if (use_pme) {
      dfree(d_tab);
      make_influence_tables();
    }
    configure_list(list_reach > 0 ? list_reach : cutoff + skin);
    forces_valid = false;
    invalidate_graph();
  }

  // Units' centres of mass scaled by s (per axis) and the box with them.
  void scale_box(const double *s) {
    scale_units<<<blocks_for(nunits, 64), 64, 0, stream>>>(nunits, d_unit_first, d_unit_size,
                                                           d_inv_mass, d_x, s[0], s[1], s[2]);
    if (nclusters)
      scale_units<<<blocks_for(nclusters, 64), 64, 0, stream>>>(
          nclusters, d_cl_first, d_cl_code, d_inv_mass, d_x, s[0], s[1], s[2]);
    HIPCHK(hipGetLastError());
    const double nb[3] = {box[0] * s[0], box[1] * s[1], box[2] * s[2]};
    set_box(nb);
  }

  double potential_energy() {
    double E[MDRT_NUM_ENERGIES];
    compute(nullptr, E);
    return potential_of(E);
  }

  // One volume move; info (or null): U0, U1, V1 - V0, W (kJ/mol, nm^3).
  bool barostat_move(double pressure_bar, double temperature, unsigned long long seed,
                     double *info) {
    const double kB = 0x1.1072c483af26dp-7;                 // kJ mol^-1 K^-1
    const double bar = 1e5 * 1e-27 * 6.02214076e23 * 1e-3;  // kJ mol^-1 nm^-3 per bar
    const double kT = kB * temperature, P = pressure_bar * bar;
    if (!(kT > 0.0)) throw HipError("the barostat needs a temperature above 0 K");
    const double v0 = box[0] * box[1] * box[2];
    if (!(mc_dv > 0.0)) mc_dv = 0.01 * v0;
    uint32_t w[4] = {(uint32_t)(uint64_t)step_count, (uint32_t)((uint64_t)step_count >> 32),
                     0u, 1u << 24};
    philox4x32_10(w, (uint32_t)seed, (uint32_t)(seed >> 32));
    auto uniform = [](uint32_t hi, uint32_t lo) {  // (0, 1), 53 bits
      return ((double)((((uint64_t)hi << 32) | lo) >> 11) + 0.5) * 0x1.0p-53;
    };
    const double dv = mc_dv * (2.0 * uniform(w[0], w[1]) - 1.0);
    const double v1 = v0 + dv;
    const double u0 = potential_energy();
    const double old_box[3] = {box[0], box[1], box[2]};
    on_device(d_save_x, d_x, 3 * (size_t)n * sizeof(double));
    const double sc = std::cbrt(v1 / v0);
    const double s3[3] = {sc, sc, sc};
    scale_box(s3);
    const double u1 = potential_energy();
    const double work = (u1 - u0) + P * dv - (double)(nunits + nclusters) * kT * log(v1 / v0);
    const bool keep = work <= 0.0 || uniform(w[2], w[3]) < exp(-work / kT);
    if (!keep) {
      on_device(d_x, d_save_x, 3 * (size_t)n * sizeof(double));
      set_box(old_box);
    }
    if (info) {
      info[0] = u0;
      info[1] = u1;
      info[2] = dv;
      info[3] = work;
    }
    ++mc_tried;
    ++mc_window_tried;
    mc_kept += keep;
    mc_window_kept += keep;
    if (mc_window_tried == 10) {
      const double rate = (double)mc_window_kept / (double)mc_window_tried;
      if (rate < 0.25) mc_dv *= 0.9;
      else if (rate > 0.75) mc_dv *= 1.1;
      mc_dv = std::min(mc_dv, 0.1 * box[0] * box[1] * box[2]);
      mc_window_tried = mc_window_kept = 0;
    }
    forces_valid = false;
    return keep;
  }

  // steps in runs of `every` steps, each followed by a volume move.
  void run_npt(int64_t steps, double dt, int graph, int every, double pressure_bar,
               double temperature, unsigned long long seed, double *kinetic) {
    if (every <= 0) throw HipError("the barostat interval must be at least one step");
    int64_t done = 0;
    while (done < steps) {
      const int64_t todo = std::min<int64_t>(every, steps - done);
      run(todo, dt, graph, nullptr);
      done += todo;
      if (todo == every) barostat_move(pressure_bar, temperature, seed, nullptr);
    }
    if (kinetic) *kinetic = kinetic_energy();
  }

  // The device's step number (read by the launches that take a first half)
  // set to the host's count before a chunk of steps.
  void put_step_count() {
    if (!(lv_friction > 0.0)) return;
    const long long h[2] = {(long long)step_count, 0};
    to_device(d_lstep, h, sizeof(h));
  }

  // ---- checkpoints ----
  // A run's state, and the positions its list was built and pruned at:
  // with those, restore_state() rebuilds the same list, so a restored run
  // takes the steps the saved one would have taken, bit for bit.  When the
  // next run would build a list anyway (positions set, a volume move), the
  // origins are the positions themselves.
  void save_state(double *x, double *v, double *forces_out, double *origin,
                  double *prune_origin, mdrtHipRunState *st) {
    if (nz != 1) throw HipError("checkpoint: not for frame batches");
    const bool kept = forces_valid && list_reach == cutoff + skin;
    if (x) to_host(x, d_x, 3 * (size_t)n * sizeof(double));
    if (v) to_host(v, d_v, 3 * (size_t)n * sizeof(double));
    if (origin) to_host(origin, kept ? d_xref : d_x, 3 * (size_t)n * sizeof(double));
    if (prune_origin)
      to_host(prune_origin, kept && d_xprune ? d_xprune : d_x, 3 * (size_t)n * sizeof(double));
    if (forces_out) {
      if (!forces_valid) evaluate(false);
      fold_pme_atoms();  // F + Fr is what the next half kick reads; folding keeps the sum
      double *d_out = dalloc<double>(3 * (size_t)n);
      forces_to_atoms<<<blocks_for(n), 256, 0, stream>>>(n, d_place_of, d_F, d_out, 0);
      to_host(forces_out, d_out, 3 * (size_t)n * sizeof(double));
      dfree(d_out);
    }
    if (st) {
      st->step_count = step_count;
      for (int k = 0; k < 3; ++k) st->box[k] = box[k];
      st->barostat_dv = mc_dv;
      st->barostat_counts[0] = mc_tried;
      st->barostat_counts[1] = mc_kept;
      st->barostat_counts[2] = mc_window_tried;
      st->barostat_counts[3] = mc_window_kept;
      st->list_kept = kept ? 1 : 0;
      st->reserved = 0;
    }
  }

  // The forces at the current positions through the list the engine holds,
  // with no build (the validity flags are clear) unless `prune` asks the
  // build chain for a prune at these positions.
  void forces_in_list(bool prune) {
    refresh_places<<<blocks_for(n), 256, 0, stream>>>(
        n, d_x, d_shift, d_q, d_place_of, use_origins() ? d_gorig : nullptr, ip.half_box[0],
        ip.half_box[1], ip.half_box[2], d_xq, d_xpa ? nullptr : d_xpq);
    HIPCHK(hipGetLastError());
    if (prune) {
      const int one = 1;
      to_device(d_flags + kFlagPrune, &one, sizeof(int));
    }
    clear_forces();
    pack_pme_atoms();
    forces(false, -1, true);
    fold_pme_atoms();
    const std::vector<int> f = read_flags();
    if (f[kFlagOverflowColumn] || f[kFlagOverflowRow] || f[kFlagBrickOverflow])
      throw HipError("restore: the list overflowed without a build");
  }

  void restore_state(const double *x, const double *v, const double *origin,
                     const double *prune_origin, const mdrtHipRunState *st) {
    if (nz != 1) throw HipError("restore: not for frame batches");
    if (!x || !v || !st) throw HipError("restore: positions, velocities and state are needed");
    if (st->box[0] != box[0] || st->box[1] != box[1] || st->box[2] != box[2])
      set_box(st->box);
    step_count = st->step_count;
    mc_dv = st->barostat_dv;
    mc_tried = st->barostat_counts[0];
    mc_kept = st->barostat_counts[1];
    mc_window_tried = st->barostat_counts[2];
    mc_window_kept = st->barostat_counts[3];
    invalidate_graph();
    const size_t bytes = 3 * (size_t)n * sizeof(double);
    to_device(d_v, v, bytes);
    if (!st->list_kept || !origin) {
      to_device(d_x, x, bytes);
      forces_valid = false;  // the next run builds at x, as the saved one would have
      return;
    }
    if (list_reach != cutoff + skin) configure_list(cutoff + skin);
    to_device(d_x, origin, bytes);
    evaluate(false);  // the list (and its prune) at the positions the run built it at
    if (d_xprune && prune_origin) {
      to_device(d_x, prune_origin, bytes);
      forces_in_list(true);  // the pruned rows where the run last pruned
    }
    to_device(d_x, x, bytes);
    forces_in_list(false);
    forces_valid = true;
  }

  // ---- force terms ----
  // Which terms a force evaluation computes (bits of MDRT_TERM_*): the
  // nonbonded pairs (LJ + direct Coulomb), the excluded and 1-4 pairs, the
  // PME reciprocal sum.  Steps always take all of them.
  int terms = 15;

  // ---- frame batches ----
  // The frame-path launches run nz frames (blockIdx.z); the per-frame
  // buffers of frame f lie f * fs bytes after frame 0's (one system: 1, 0).
  int nz = 1;
  size_t fs = 0;
  char *d_arena = nullptr;             // nz slabs of fs bytes
  std::vector<size_t> frame_bytes;     // per-frame size of each buffer
  std::vector<void **> frame_extra;    // buffers the frame evaluator adds

  // Every buffer a frame's build, pair loop and PME chain own.
  std::vector<void **> frame_buffers() {
    std::vector<void **> v = {
        (void **)&d_x, (void **)&d_xref, (void **)&d_shift, (void **)&d_place_of,
        (void **)&d_atom_at, (void **)&d_type_at, (void **)&d_xq, (void **)&d_gorig,
        (void **)&d_xpq,
        (void **)&d_F, (void **)&d_col_count, (void **)&d_bucket, (void **)&d_bb_lo,
        (void **)&d_bb_hi, (void **)&d_cluster_col, (void **)&d_real_mask,
        (void **)&d_col_cl_start, (void **)&d_col_cl_count, (void **)&d_counts,
        (void **)&d_rows, (void **)&d_row_count, (void **)&d_ex, (void **)&d_epair,
        (void **)&d_flags};
    if (nspecial) v.push_back((void **)&d_espec);
    if (use_pme) {
      v.push_back((void **)&d_cplx);
      v.push_back((void **)&d_erecip);
      v.push_back((void **)&bp.count);
      v.push_back((void **)&bp.list);
    }
    for (void **p : frame_extra) v.push_back(p);
    return v;
  }

  // Moves the per-frame buffers into one arena of `frames` slabs, each slab
  // starting as a copy of the single system's buffers.
  void batch(int frames) {
    if (nz != 1 || d_arena) throw HipError("already batched");
    if (ntuples || nclusters)
      throw HipError("frame batches take systems without bonded tuples or constraint clusters");
    if (frames < 1) throw HipError("a batch needs at least one frame");
    if (use_pme && !plane_fft())
      throw HipError("frame batches need a PME grid whose xy plane the plane transform holds");
    if (prune_skin > 0) throw HipError("frame batches build their lists without pruning");
    std::vector<void **> bufs = frame_buffers();
    std::vector<size_t> off(bufs.size());
    frame_bytes.assign(bufs.size(), 0);
    size_t total = 0;
    for (size_t i = 0; i < bufs.size(); ++i) {
      void *base = nullptr;
      size_t size = 0;
      HIPCHK(hipMemGetAddressRange(&base, &size, *bufs[i]));
      if (base != *bufs[i]) throw HipError("a per-frame buffer is not an allocation");
      frame_bytes[i] = size;
      off[i] = total;
      total += (size + 255) & ~(size_t)255;
    }
    char *arena = dalloc<char>((size_t)frames * total);
    for (int f = 0; f < frames; ++f)
      for (size_t i = 0; i < bufs.size(); ++i)
        on_device(arena + (size_t)f * total + off[i], *bufs[i], frame_bytes[i]);
    HIPCHK(hipStreamSynchronize(stream));
    for (size_t i = 0; i < bufs.size(); ++i) {
      dfree(*bufs[i]);
      *bufs[i] = arena + off[i];
    }
    d_arena = arena;
    nz = frames;
    fs = total;
    invalidate_graph();
  }

  // Back to one system: the buffers of frame 0, each in its own allocation.
  void unbatch() {
    if (!d_arena) return;
    std::vector<void **> bufs = frame_buffers();
    for (size_t i = 0; i < bufs.size(); ++i) {
      char *p = dalloc<char>(frame_bytes[i]);
      on_device(p, *bufs[i], frame_bytes[i]);
      *bufs[i] = p;
    }
    HIPCHK(hipStreamSynchronize(stream));
    dfree(d_arena);
    nz = 1;
    fs = 0;
    invalidate_graph();
  }

  // Captured step sequences, one per (steps, dt, measured): a run of K
  // steps in chunks of G replays the G-step graph and, for a remainder, a
  // graph of K mod G steps, so no run falls back to launch-by-launch steps.
  struct StepGraph {
    int steps;
    double dt;
    bool measured;
    hipGraphExec_t exec;
  };
  static constexpr size_t kMaxStepGraphs = 4;
  std::vector<StepGraph> step_graphs;
  uint64_t graph_generation = 0;  // bumped whenever captured step graphs go stale
  void invalidate_graph() {
    for (StepGraph &g_ : step_graphs) hipGraphExecDestroy(g_.exec);
    step_graphs.clear();
    ++graph_generation;
  }
  hipGraphExec_t find_graph(int steps, double dt, bool measured) const {
    for (const StepGraph &g_ : step_graphs)
      if (g_.steps == steps && g_.dt == dt && g_.measured == measured) return g_.exec;
    return nullptr;
  }

  bool forces_valid = false;
  long long builds_seen = 0;

  ~Engine() { release(); }

  // Transfers ordered on the engine's stream and complete on return.
  void to_host(void *dst, const void *src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  void to_device(void *dst, const void *src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  void on_device(void *dst, const void *src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream));
  }

  void release() {
    if (d_arena) {  // per-frame pointers lie inside the arena
      for (void **b : frame_buffers()) *b = nullptr;
      dfree(d_arena);
      nz = 1;
      fs = 0;
    }
    invalidate_graph();
    dfree(d_timing);
    timing_slots = 0;
    dfree(d_x); dfree(d_v); dfree(d_xref); dfree(d_shift); dfree(d_q);
    dfree(d_inv_mass); dfree(d_type); dfree(d_place_of); dfree(d_save_x); dfree(d_save_v);
    dfree(d_xprune); dfree(d_xpa); dfree(d_Fr); dfree(d_pme_n);
    release_list();
    dfree(d_excl_start); dfree(d_excl_atoms); dfree(d_flags); dfree(d_lj);
    dfree(d_tuple_atoms); dfree(d_tuple_kind); dfree(d_tuple_param); dfree(d_etup);
    dfree(d_clen); dfree(d_cl_first); dfree(d_cl_code);
    dfree(d_special); dfree(d_is14); dfree(d_isexcl); dfree(d_p14_sigma);
    dfree(d_p14_eps); dfree(d_p14_qq); dfree(d_espec);
    fft.release();
    dfree(d_cplx); dfree(d_tab); dfree(d_erecip);
    dfree(d_dft1); dfree(d_dft2);
    dfree(bp.count); dfree(bp.list);
    dfree(d_unit_first); dfree(d_unit_size); dfree(d_lstep);
    if (ev_fork) hipEventDestroy(ev_fork);
    if (ev_join) hipEventDestroy(ev_join);
    ev_fork = ev_join = nullptr;
    if (side) hipStreamDestroy(side);
    side = nullptr;
    if (stream) hipStreamDestroy(stream);
    stream = nullptr;
  }

  void release_list() {
    dfree(d_atom_at); dfree(d_type_at); dfree(d_xq); dfree(d_gorig); dfree(d_xpq); dfree(d_F); dfree(d_col_count);
    dfree(d_bucket); dfree(d_bb_lo); dfree(d_bb_hi); dfree(d_cluster_col);
    dfree(d_real_mask); dfree(d_col_cl_start); dfree(d_col_cl_count); dfree(d_counts);
    dfree(d_rows); dfree(d_row_count); dfree(d_epair); dfree(d_ex);
    dfree(d_rows_in); dfree(d_row_count_in);
  }

  // Column grid, capacities and the list buffers for a search reach.
  void configure_list(double reach, int col_cap_min = 0, int row_cap_min = 0,
                      int ex_cap_min = 0) {
    const double volume = box[0] * box[1] * box[2];
    const double rho = n / volume;
    const double target = std::cbrt((double)kCluster / rho);
    GridParams ng{};
    for (int k = 0; k < 3; ++k) {
      ng.box[k] = box[k];
      ng.inv_box[k] = 1.0 / box[k];
      ng.half_box[k] = 0.5 * box[k];
    }
    ng.ncx = std::max(1, (int)floor(box[0] / target));
    ng.ncy = std::max(1, (int)floor(box[1] / target));
    ng.ncol = ng.ncx * ng.ncy;
    ng.wx = box[0] / ng.ncx;
    ng.wy = box[1] / ng.ncy;
    const double per_col = (double)n / ng.ncol;
    int cap = (int)(2.0 * per_col + 128.0);
    cap = std::max(cap, col_cap_min);
    cap = (cap + 63) & ~63;
    if (cap > kSortCapMax) cap = kSortCapMax;
    ng.col_cap = cap;
    const double rb = reach + 1e-4;
    ng.rlist2 = (float)(rb * rb);
    ng.rx = (int)(rb / ng.wx) + 1;
    ng.ry = (int)(rb / ng.wy) + 1;
    // every column is padded to whole groups of kGroup clusters
    const int gp = kCluster * kGroup;
    const int places = n + (gp - 1) * ng.ncol;
    const int ngroups_max = (places + gp - 1) / gp;
    ng.max_clusters = ngroups_max * kGroup;
    // entries of a group row: clusters whose boxes are within the reach of
    // the group's box (all of them: a bound on the half list's share)
    const double h = (double)kCluster / (rho * ng.wx * ng.wy);
    double est = (ng.wx + 2 * rb) * (ng.wy + 2 * rb) *
                 (std::min(kGroup * h, box[2]) + 2 * rb) / (ng.wx * ng.wy * h);
    est = std::min(est, (double)ng.max_clusters * 3.0);
    int rc_ = (int)(1.25 * est + 64);
    rc_ = std::max(rc_, row_cap_min);
    ng.row_cap = (rc_ + kSegmentMax - 1) / kSegmentMax * kSegmentMax;
    ng.max_excl = 16;
    const int nex = std::max(ex_cap_min, ngroups_max * 16 + 1024);
    const bool same = d_rows && ng.ncol == g.ncol && ng.col_cap == g.col_cap &&
                      ng.row_cap == g.row_cap && ng.max_clusters == g.max_clusters &&
                      nex == ex_cap && seg_len == seg_len_cfg;
    g = ng;
    list_reach = reach;
    if (same) return;
    max_groups = ngroups_max;
    nseg = g.row_cap / seg_len;
    seg_len_cfg = seg_len;
    ex_cap = nex;
    release_list();
    max_places = g.max_clusters * kCluster;
    d_atom_at = dalloc<int>(max_places);
    HIPCHK(hipMemset(d_atom_at, 0xFF, max_places * sizeof(int)));  // -1: no atom
    HIPCHK(hipDeviceSynchronize());
    d_type_at = dalloc<int>(max_places);
    d_xq = dalloc<float4>(max_places);
    d_gorig = dalloc<float4>(max_places / (kCluster * kGroup) + 1);
    d_xpq = dalloc<double4>(max_places);
    d_F = dalloc<long long>(3 * (size_t)max_places);
    d_col_count = dalloc<int>(g.ncol);
    d_bucket = dalloc<int>((size_t)g.ncol * g.col_cap);
    d_bb_lo = dalloc<float4>(g.max_clusters);
    d_bb_hi = dalloc<float4>(g.max_clusters);
    d_cluster_col = dalloc<int>(g.max_clusters);
    d_real_mask = dalloc<uint32_t>(g.max_clusters);
    d_col_cl_start = dalloc<int>(g.ncol);
    d_col_cl_count = dalloc<int>(g.ncol);
    d_counts = dalloc<int>(4);
    d_rows = dalloc<PairEntry>((size_t)max_groups * g.row_cap);
    d_row_count = dalloc<int>(max_groups);
    d_ex = dalloc<uint64_t>((size_t)ex_cap * kGroup);
    if (prune_skin > 0) {
      d_rows_in = dalloc<PairEntry>((size_t)max_groups * g.row_cap);
      d_row_count_in = dalloc<int>(max_groups);
    }
    d_epair = dalloc<double>(2 * (size_t)max_groups * nseg);
    forces_valid = false;
    invalidate_graph();
  }

  void create(const mdrtHipSystemDesc *d) {
    n = (int)d->n;
    if (n <= 0) throw HipError("a system needs particles");
    // Pair-row segments: 16 entries give a box of tens of thousands of atoms
    // waves enough to fill the GPU; from about 200,000 atoms the groups
    // alone do, and 48-entry segments save per-wave set-up (11 % per step
    // measured at 999,000 atoms).  The launch-shape trial may change it.
    if (n >= 200000) seg_len = 48;
    ntypes = d->ntypes;
    if (ntypes < 1 || ntypes > kMaxTypes)
      throw HipError("ntypes must be in [1, " + std::to_string(kMaxTypes) + "]");
    for (int k = 0; k < 3; ++k) box[k] = d->box[k];
    cutoff = d->cutoff;
    switch_from = d->switch_from;
    skin = d->skin > 0 ? d->skin : 0.1;
    prune_skin = d->prune_skin > 0 ? d->prune_skin : 0.0;
    if (prune_skin >= skin) prune_skin = 0.0;  // nothing to prune
    beta = d->beta;
    lj_modifier = d->lj_modifier;
    coulomb_shift = d->coulomb_shift;
    dispersion = d->dispersion;
    pme_order = d->pme_order > 0 ? d->pme_order : 4;
    for (int k = 0; k < 3; ++k) pme_grid[k] = d->pme_grid[k];
    const double rl = cutoff + skin;
    for (int k = 0; k < 3; ++k)
      if (box[k] <= 2.0 * rl)
        throw HipError("every box edge must exceed 2 (cutoff + skin)");
    if (lj_modifier < 0 || lj_modifier > 3) throw HipError("unknown Lennard-Jones modifier");
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    pstream = side;
    lstream = stream;

    h_q.assign(d->charges, d->charges + n);
    h_mass.assign(d->masses, d->masses + n);
    h_type.assign(d->types, d->types + n);
    h_sigma.assign(d->sigma, d->sigma + ntypes * ntypes);
    h_eps.assign(d->epsilon, d->epsilon + ntypes * ntypes);
    for (int t : h_type)
      if (t < 0 || t >= ntypes) throw HipError("type index out of range");

    d_x = dalloc<double>(3 * (size_t)n);
    d_v = dalloc<double>(3 * (size_t)n);
    d_xref = dalloc<double>(3 * (size_t)n);
    d_shift = dalloc<double>(3 * (size_t)n);
    d_save_x = dalloc<double>(3 * (size_t)n);
    d_save_v = dalloc<double>(3 * (size_t)n);
    HIPCHK(hipMemcpy(d_x, d->positions, 3 * n * sizeof(double), hipMemcpyHostToDevice));
    if (d->velocities)
      HIPCHK(hipMemcpy(d_v, d->velocities, 3 * n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipDeviceSynchronize());
    d_q = dupload(h_q);
    std::vector<double> im(n);
    for (int i = 0; i < n; ++i) im[i] = h_mass[i] > 0 ? 1.0 / h_mass[i] : 0.0;
    d_inv_mass = dupload(im);
    d_type = dupload(h_type);
    d_place_of = dalloc<int>(n);
    d_flags = dalloc<int>(kNumFlags);

    // exclusions: symmetric CSR and the i<j list
    std::vector<std::vector<int>> rows(n);
    for (int64_t e = 0; e < d->nexclusions; ++e) {
      int i = d->exclusions[2 * e], j = d->exclusions[2 * e + 1];
      if (i < 0 || j < 0 || i >= n || j >= n || i == j) throw HipError("bad exclusion");
      if (i > j) std::swap(i, j);
      rows[i].push_back(j);
      rows[j].push_back(i);
      h_excl_pairs.push_back(i);
      h_excl_pairs.push_back(j);
    }
    std::vector<int> es(n + 1, 0), ea;
    for (int i = 0; i < n; ++i) {
      std::sort(rows[i].begin(), rows[i].end());
      if ((int)rows[i].size() > 16) throw HipError("more than 16 exclusions of one particle");
      es[i + 1] = es[i] + (int)rows[i].size();
      for (int j : rows[i]) ea.push_back(j);
    }
    d_excl_start = dupload(es);
    d_excl_atoms = dupload(ea);

    upload_lj();

    // nonbonded constants
    for (int k = 0; k < 3; ++k) {
      nbp.box[k] = (float)box[k];
      nbp.box_lo[k] = (float)(box[k] - (double)nbp.box[k]);
      nbp.half_box[k] = (float)(0.5 * box[k]);
    }
    nbp.rc = (float)cutoff;
    nbp.rc2 = (float)(cutoff * cutoff);
    nbp.rs = (float)switch_from;
    nbp.beta = (float)beta;
    nbp.mb2l2e = (float)(-beta * beta / M_LN2);
    nbp.tbs = (float)(2.0 * beta / sqrt(M_PI));
    nbp.coulomb = (float)kCoulomb;
    nbp.shift_c = (coulomb_shift && beta > 0) ? (float)(erfc(beta * cutoff) / cutoff) : 0.0f;
    nbp.ntypes = ntypes;
    nbp.lj_modifier = lj_modifier;
    nbp.shift12 = (float)(pow(cutoff, -12.0) / 12.0);
    nbp.shift6 = (float)(pow(cutoff, -6.0) / 6.0);
    if (lj_modifier == kLJForceSwitch) {
      const double rc = cutoff, rs = switch_from, w = rc - rs;
      if (!(w > 0)) throw HipError("force switch needs switch_from < cutoff");
      double A[2], Bc[2], Cc[2];
      const double ns[2] = {12.0, 6.0};
      for (int k = 0; k < 2; ++k) {
        const double nn = ns[k];
        A[k] = -nn * ((nn + 4.0) * rc - (nn + 1.0) * rs) / (pow(rc, nn + 2.0) * w * w);
        Bc[k] = nn * ((nn + 3.0) * rc - (nn + 1.0) * rs) / (pow(rc, nn + 2.0) * w * w * w);
        Cc[k] = pow(rc, -nn) - A[k] / 3.0 * pow(w, 3.0) - Bc[k] / 4.0 * pow(w, 4.0);
      }
      nbp.sw_a12 = (float)(A[0] / 12.0); nbp.sw_b12 = (float)(Bc[0] / 12.0);
      nbp.sw_c12 = (float)(Cc[0] / 12.0);
      nbp.sw_a6 = (float)(A[1] / 6.0); nbp.sw_b6 = (float)(Bc[1] / 6.0);
      nbp.sw_c6 = (float)(Cc[1] / 6.0);
    }
    if (lj_modifier == kLJPotentialSwitch && !(switch_from < cutoff))
      throw HipError("potential switch needs switch_from < cutoff");

    // special pairs: exclusions (Ewald correction) and 1-4 pairs
    {
      std::vector<int2> sp;
      std::vector<int> is14, isex;
      for (size_t e = 0; e < h_excl_pairs.size() / 2; ++e) {
        sp.push_back(make_int2(h_excl_pairs[2 * e], h_excl_pairs[2 * e + 1]));
        is14.push_back(-1);
        isex.push_back(1);
      }
      std::unordered_map<long long, size_t> where;
      where.reserve(sp.size() * 2 + 1);
      for (size_t e = 0; e < sp.size(); ++e) {
        const int a = std::min(sp[e].x, sp[e].y), b = std::max(sp[e].x, sp[e].y);
        where.emplace(((long long)a << 32) | (unsigned)b, e);
      }
      for (int64_t p = 0; p < d->npairs14; ++p) {
        int i = d->pairs14[2 * p], j = d->pairs14[2 * p + 1];
        if (i > j) std::swap(i, j);
        const auto it = where.find(((long long)i << 32) | (unsigned)j);
        if (it == where.end())
          throw HipError("a 1-4 pair must also be an excluded pair");
        is14[it->second] = (int)p;
      }
      nspecial = (int)sp.size();
      d_special = dupload(sp);
      d_is14 = dupload(is14);
      d_isexcl = dupload(isex);
      std::vector<double> s14(d->npairs14), e14(d->npairs14), q14(d->npairs14);
      for (int64_t p = 0; p < d->npairs14; ++p) {
        s14[p] = d->pairs14_sigma[p];
        e14[p] = d->pairs14_eps[p];
        q14[p] = d->pairs14_qq[p];
      }
      d_p14_sigma = dupload(s14);
      d_p14_eps = dupload(e14);
      d_p14_qq = dupload(q14);
      for (int k = 0; k < 3; ++k) spp.box[k] = box[k];
      spp.beta = beta;
      spp.coulomb = kCoulomb;
      spp.rc = cutoff;
      spp.rs = switch_from;
      spp.lj_modifier = lj_modifier;
      d_espec = dalloc<double>(3 * (size_t)std::max(1, blocks_for(nspecial)));
    }

    // bonded tuples, sorted by kind so that a wave mostly runs one kind
    if (d->ntuples > 0) {
      if (!d->tuple_atoms || !d->tuple_kind || !d->tuple_param)
        throw HipError("tuples without their atoms, kinds or parameters");
      std::vector<int> order((size_t)d->ntuples);
      for (int64_t t = 0; t < d->ntuples; ++t) {
        const int k = d->tuple_kind[t];
        if (k < 0 || k >= kTupleKinds) throw HipError("a tuple kind outside MDRT_TUPLE_*");
        const int members = k <= kTupleUreyBradley ? 2 : (k == kTupleAngle ? 3 : 4);
        for (int m = 0; m < members; ++m) {
          const int a = d->tuple_atoms[4 * t + m];
          if (a < 0 || a >= n) throw HipError("a tuple atom out of range");
        }
        order[t] = (int)t;
      }
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return d->tuple_kind[a] < d->tuple_kind[b];
      });
      std::vector<int4> ta;
      std::vector<int> tk;
      std::vector<double4> tp;
      for (int t : order) {
        const int32_t *a = d->tuple_atoms + 4 * (size_t)t;
        const double *q = d->tuple_param + 4 * (size_t)t;
        const int k = d->tuple_kind[t];
        const int members = k <= kTupleUreyBradley ? 2 : (k == kTupleAngle ? 3 : 4);
        ta.push_back(make_int4(a[0], a[1], members > 2 ? a[2] : 0, members > 3 ? a[3] : 0));
        tk.push_back(k);
        tp.push_back(make_double4(q[0], q[1], q[2], q[3]));
      }
      ntuples = (int)d->ntuples;
      d_tuple_atoms = dupload(ta);
      d_tuple_kind = dupload(tk);
      d_tuple_param = dupload(tp);
      d_etup = dalloc<double>((size_t)kTupleKinds * kTupleBlocksMax);
    }

    host_constants();

    // PME
    use_pme = beta > 0;
    if (use_pme && pme_atoms) {
      d_xpa = dalloc<double4>(n);
      d_Fr = dalloc<long long>(3 * (size_t)n);
      HIPCHK(hipMemset(d_Fr, 0, 3 * (size_t)n * sizeof(long long)));
      const int cnt[4] = {n, 0, 0, 0};
      d_pme_n = dalloc<int>(4);
      HIPCHK(hipMemcpy(d_pme_n, cnt, sizeof(cnt), hipMemcpyHostToDevice));
    }
    if (use_pme) {
      for (int k = 0; k < 3; ++k) {
        if (pme_grid[k] < 2 * pme_order) throw HipError("PME grid smaller than twice the order");
        if (pme_grid[k] > 1024) throw HipError("PME grid edge above 1024");
        pmp.K[k] = pme_grid[k];
        pmp.box[k] = box[k];
        pmp.inv_box[k] = 1.0 / box[k];
      }
      if (pme_order != 4 && pme_order != 6 && pme_order != 8)
        throw HipError("PME order must be 4, 6 or 8");
      pmp.order = pme_order;
      fft.make(pme_grid);
      const size_t nh = (size_t)pme_grid[0] * pme_grid[1] * (pme_grid[2] / 2 + 1);
      d_cplx = dalloc<cpx<float>>(nh);
      make_influence_tables();
      d_erecip = dalloc<double>(std::max(fft.x_blocks(), fft.H()));
    }

    // integration units: waters, then the other particles
    {
      std::vector<int> uf, us;
      std::vector<char> in_water(n, 0);
      for (int64_t w = 0; w < d->nwaters; ++w) {
        const int o = d->water_first[w];
        if (o < 0 || o + 2 >= n) throw HipError("bad water index");
        if (in_water[o] || in_water[o + 1] || in_water[o + 2])
          throw HipError("two waters share an atom");
        uf.push_back(o);
        us.push_back(3);
        in_water[o] = in_water[o + 1] = in_water[o + 2] = 1;
      }
      // constraint clusters: the atoms first .. first + size - 1
      std::vector<double> clen(n, 0.0);
      std::vector<int> cf, cc;
      for (int64_t c = 0; c < d->nclusters; ++c) {
        const int first = d->cluster_first[c], size = d->cluster_size[c];
        const int centre = d->cluster_center[c];
        if (size < 2 || size > 5) throw HipError("a constraint cluster of other than 2-5 atoms");
        if (first < 0 || first + size > n || centre < 0 || centre >= size)
          throw HipError("a constraint cluster out of range");
        for (int a = first; a < first + size; ++a) {
          if (in_water[a]) throw HipError("a constraint cluster overlaps a water or a cluster");
          in_water[a] = 1;
          if (a != first + centre) {
            clen[a] = d->cluster_length ? d->cluster_length[a] : 0.0;
            if (!(clen[a] > 0.0)) throw HipError("a constraint length must be positive");
          }
          if (!(h_mass[a] > 0.0)) throw HipError("a constrained atom needs a mass");
        }
        cf.push_back(first);
        cc.push_back(size | kUnitCluster | (centre << kUnitCentreShift));
      }
      nclusters = (int)d->nclusters;
      if (nclusters) {
        d_clen = dupload(clen);
        d_cl_first = dupload(cf);
        d_cl_code = dupload(cc);
      }
      ip.clen = d_clen;
      ip.inv_mass = d_inv_mass;
      for (int i = 0; i < n; ++i)
        if (!in_water[i]) {
          uf.push_back(i);
          us.push_back(1);
        }
      nunits = (int)uf.size();
      d_unit_first = dupload(uf);
      d_unit_size = dupload(us);
      for (int k = 0; k < 3; ++k) ip.half_box[k] = 0.5 * box[k];
      ip.skin_half2 = 0.25 * skin * skin;
      if (prune_skin > 0) {
        d_xprune = dalloc<double>(3 * (size_t)n);
        ip.prune_half2 = 0.25 * prune_skin * prune_skin;
        ip.x_prune = d_xprune;
      } else {
        ip.prune_half2 = 0.0;
        ip.x_prune = nullptr;
      }
      if (d->nwaters > 0) {
        const int o = d->water_first[0];
        ip.m_o = h_mass[o];
        ip.m_h = h_mass[o + 1];
        ip.d_oh = d->d_oh;
        ip.d_hh = d->d_hh;
        const double total = ip.m_o + 2 * ip.m_h;
        ip.rc = 0.5 * ip.d_hh;
        const double height = sqrt(ip.d_oh * ip.d_oh - ip.rc * ip.rc);
        ip.ra = 2.0 * ip.m_h * height / total;
        ip.rb = height - ip.ra;
      }
    }
    configure_list(cutoff + skin);
    if (use_pme) make_bricks();
    HIPCHK(hipStreamSynchronize(stream));
  }

  // Lennard-Jones table (6 c6, 12 c12) with c6 = 4 eps sigma^6, c12 = 4 eps
  // sigma^12: the factors of the force, so that it needs no scaling.  Written
  // in place once allocated (captured graphs hold the pointer).
  void upload_lj() {
    std::vector<float2> lj(ntypes * ntypes);
    for (int k = 0; k < ntypes * ntypes; ++k) {
      const double s6 = pow(h_sigma[k], 6.0);
      lj[k] = make_float2((float)(6.0 * 4.0 * h_eps[k] * s6),
                          (float)(12.0 * 4.0 * h_eps[k] * s6 * s6));
    }
    if (!d_lj) d_lj = dalloc<float2>(lj.size());
    to_device(d_lj, lj.data(), lj.size() * sizeof(float2));
  }

  // Host constants: the Ewald self and net-charge terms, the dispersion
  // correction.
  void host_constants() {
    {
      double sq = 0, net = 0;
      for (double qq : h_q) {
        sq += qq * qq;
        net += qq;
      }
      const double volume = box[0] * box[1] * box[2];
      e_self = beta > 0 ? -kCoulomb * beta / sqrt(M_PI) * sq -
                              kCoulomb * M_PI * net * net / (2.0 * volume * beta * beta)
                        : 0.0;
      e_disp = 0;
      if (dispersion) {
        std::vector<double> num(ntypes, 0.0);
        for (int t : h_type) num[t] += 1.0;
        double sum = 0;
        for (int a = 0; a < ntypes; ++a)
          for (int b = 0; b < ntypes; ++b)
            sum += num[a] * (num[b] - (a == b ? 1.0 : 0.0)) * 4.0 * h_eps[a * ntypes + b] *
                   pow(h_sigma[a * ntypes + b], 6.0);
        const size_t ne = h_excl_pairs.size() / 2;
        for (size_t e = 0; e < ne; ++e) {
          const int a = h_type[h_excl_pairs[2 * e]], b = h_type[h_excl_pairs[2 * e + 1]];
          sum -= 2.0 * 4.0 * h_eps[a * ntypes + b] * pow(h_sigma[a * ntypes + b], 6.0);
        }
        const double nn = n, pairs = nn * (nn - 1.0) - 2.0 * (double)ne;
        const double mean = pairs > 0 ? sum / pairs : 0.0;
        e_disp = -2.0 * M_PI / (3.0 * volume) * nn * nn * mean / (cutoff * cutoff * cutoff);
      }
    }
  }

  void make_influence_tables() {
    const int Kmax = std::max({pme_grid[0], pme_grid[1], pme_grid[2]});
    std::vector<float> tab(6 * (size_t)Kmax, 0.0f);
    const double volume = box[0] * box[1] * box[2];
    for (int a = 0; a < 3; ++a) {
      const int K = pme_grid[a];
      // |b(k)|^2 moduli (Influence.cpp:168-191)
      std::vector<double> m(pme_order + 1, 0.0), next(pme_order + 1, 0.0);
      m[1] = 1.0;
      for (int nn = 3; nn <= pme_order; ++nn) {
        std::fill(next.begin(), next.end(), 0.0);
        for (int xx = 1; xx < nn; ++xx) next[xx] = (xx * m[xx] + (nn - xx) * m[xx - 1]) / (nn - 1);
        m.swap(next);
      }
      for (int k = 0; k < K; ++k) {
        double re = 0, im = 0;
        for (int j = 0; j <= pme_order - 2; ++j) {
          const double ang = 2.0 * M_PI * (double)k * j / (double)K;
          re += m[j + 1] * cos(ang);
          im += m[j + 1] * sin(ang);
        }
        const double modulus = 1.0 / (re * re + im * im);
        const int ks = k <= K / 2 ? k : k - K;
        const double mk = (double)ks / box[a];
        double t = exp(-M_PI * M_PI * mk * mk / (beta * beta)) * modulus;
        if (a == 0) t *= kCoulomb / (M_PI * volume);
        tab[a * Kmax + k] = (float)t;
        tab[(3 + a) * Kmax + k] = (float)mk;
      }
    }
    d_tab = dupload(tab);
    for (int a = 0; a < 3; ++a) {
      ct.t[a] = d_tab + a * Kmax;
      ct.m[a] = d_tab + (3 + a) * Kmax;
    }
    ct.K3 = pme_grid[2];
    ct.pb2 = (float)(M_PI * M_PI / (beta * beta));
  }

  // ---- launches ----

  void request_build() { nb_request_build<<<dim3(1, 1, nz), 64, 0, stream>>>(d_flags, fs); }

  void launch_bin() {
    nb_bin<<<dim3(blocks_for(n), 1, nz), 256, 0, lstream>>>(n, d_x, d_shift, d_xref,
                                                          d_col_count, d_bucket, d_flags, g, fs);
  }
  void launch_sort() {
    nb_sort<<<dim3(g.ncol, 1, nz), 256, 0, lstream>>>(
        d_x, d_shift, d_q, d_type, d_col_count, d_bucket, d_atom_at, d_place_of, d_xq, d_gorig,
        d_xpq, d_type_at, d_bb_lo, d_bb_hi, d_cluster_col, d_real_mask, d_col_cl_start,
        d_col_cl_count, d_counts, d_flags, g, use_origins() ? 1 : 0, fs);
  }
  void launch_search() {
    nb_search<<<dim3(max_groups, 1, nz), 256, 0, lstream>>>(
        d_bb_lo, d_bb_hi, d_cluster_col, d_real_mask, d_col_cl_start, d_col_cl_count,
        d_atom_at, d_place_of, d_excl_start, d_excl_atoms, d_counts, d_col_count, d_rows,
        d_row_count, d_ex, ex_cap, d_flags, g, fs);
  }
  void launch_prune() {
    if (prune_skin <= 0) return;
    if ((size_t)g.row_cap * sizeof(int) > 64 * 1024)
      throw HipError("pruned list: rows longer than 16,384 entries");
    const double reach = cutoff + prune_skin + 1e-4;
    nb_prune<<<max_groups, 512, (size_t)g.row_cap * sizeof(int), lstream>>>(
        d_xq, d_gorig, d_rows, d_row_count, d_rows_in, d_row_count_in, d_counts, d_atom_at,
        d_x, d_xprune, g.row_cap, (float)(reach * reach), nbp.box[0], nbp.box[1], nbp.box[2],
        nbp.box_lo[0], nbp.box_lo[1], nbp.box_lo[2], d_flags);
  }
  void build() {
    launch_bin();
    launch_sort();
    launch_search();
    launch_prune();
  }

  // One launch: plane y = 0 of the grid takes the excluded and 1-4 pairs,
  // the planes above the segments of the group rows.
  template <int MOD, bool COUL, bool ENERGY, bool ORG>
  void launch_pairs_t() {
    SpecialArgs sa{};
    sa.npairs = (terms & 2) ? nspecial : 0;
    sa.pairs = d_special;
    sa.is14 = d_is14;
    sa.p14_sigma = d_p14_sigma;
    sa.p14_eps = d_p14_eps;
    sa.p14_qq = d_p14_qq;
    sa.x = d_x;
    sa.q = d_q;
    sa.place_of = d_place_of;
    sa.energy_partial = ENERGY ? d_espec : nullptr;
    sa.p = spp;
    if (const int tb = tuple_blocks()) {
      TupleArgs ta{};
      ta.ntuples = ntuples;
      ta.atoms = d_tuple_atoms;
      ta.kind = d_tuple_kind;
      ta.param = d_tuple_param;
      ta.energy_partial = ENERGY ? d_etup : nullptr;
      bonded_tuples<<<tb, 256, 0, lstream>>>(ta, d_x, d_place_of, box[0], box[1], box[2], d_F);
    }
    const dim3 grid(std::max(blocks_for(max_groups, 4), blocks_for(nspecial)),
                    (terms & 1) ? nseg + 1 : 1, nz);
    const bool pruned = prune_skin > 0;
    const size_t lds = ntypes * ntypes * sizeof(float2);
    if constexpr (ORG)
      nb_pairs_org<MOD, COUL, ENERGY><<<grid, 256, lds, lstream>>>(
          d_xq, d_gorig, d_type_at, d_real_mask, pruned ? d_rows_in : d_rows,
          pruned ? d_row_count_in : d_row_count, d_ex, g.row_cap, seg_len, max_groups, d_counts,
          d_lj, nbp, d_F, d_epair, d_flags, sa, fs);
    else if (nz > 1)  // a batch of frames
      nb_pairs_batch<MOD, COUL, ENERGY><<<grid, 256, lds, lstream>>>(
          d_xq, d_gorig, d_type_at, d_real_mask, pruned ? d_rows_in : d_rows,
          pruned ? d_row_count_in : d_row_count, d_ex, g.row_cap, seg_len, max_groups, d_counts,
          d_lj, nbp, d_F, d_epair, d_flags, sa, fs);
    else
      nb_pairs<MOD, COUL, ENERGY, false><<<grid, 256, lds, lstream>>>(
          d_xq, d_gorig, d_type_at, d_real_mask, pruned ? d_rows_in : d_rows,
          pruned ? d_row_count_in : d_row_count, d_ex, g.row_cap, seg_len, max_groups, d_counts,
          d_lj, nbp, d_F, d_epair, d_flags, sa, fs);
  }

  template <int MOD, bool ORG>
  void launch_pairs_o(bool energy) {
    const bool coul = beta > 0;
    if (coul && energy) launch_pairs_t<MOD, true, true, ORG>();
    else if (coul) launch_pairs_t<MOD, true, false, ORG>();
    else if (energy) launch_pairs_t<MOD, false, true, ORG>();
    else launch_pairs_t<MOD, false, false, ORG>();
  }
  template <int MOD>
  void launch_pairs(bool energy) {
    if (use_origins()) launch_pairs_o<MOD, true>(energy);
    else launch_pairs_o<MOD, false>(energy);
  }

  void pairs(bool energy) {
    switch (lj_modifier) {
      case kLJNone: launch_pairs<kLJNone>(energy); break;
      case kLJPotentialShift: launch_pairs<kLJPotentialShift>(energy); break;
      case kLJForceSwitch: launch_pairs<kLJForceSwitch>(energy); break;
      default: launch_pairs<kLJPotentialSwitch>(energy); break;
    }
  }


  // ---- bricks of the spread and gather ----
  BrickParams bp{};
  int brick_tl = 8;            // complex lines of a brick's z transform
  int brick_tl_max = 8;        // the widest the shape may take (autotune)
  int nbricks = 0;

  // Brick shape from the z length (LDS of 2 TL K3 64-bit accumulators, two
  // TL K3 complex buffers and the K3 twiddles within 60 KB) and the
  // per-brick atom lists,
  // sized from the mean count with room for fluctuations (grown on
  // overflow).
  void make_bricks(int cap_min = 0) {
    const int K2 = pme_grid[2];
    brick_tl = brick_tl_max;
    while (brick_tl > 1 && (size_t)(32 * brick_tl + 8) * K2 > 60 * 1024) brick_tl /= 2;
    if ((size_t)(32 * brick_tl + 8) * K2 > 60 * 1024) throw HipError("PME grid too long along z");
    const int R = 2 * brick_tl, BY = R >= 8 ? 4 : 2, BX = R / BY;
    bp.nbx = (pme_grid[0] + BX - 1) / BX;
    bp.nby = (pme_grid[1] + BY - 1) / BY;
    nbricks = bp.nbx * bp.nby;
    const double touched = ((pme_order - 1.0) / BX + 1.0) * ((pme_order - 1.0) / BY + 1.0);
    const double mean = n * touched / nbricks;
    int cap = (int)(1.5 * mean + 128.0);
    cap = std::max(cap, cap_min);
    cap = (cap + 63) & ~63;
    invalidate_graph();  // captured graphs point at the old lists
    dfree(bp.count);
    dfree(bp.list);
    bp.cap = cap;
    bp.count = dalloc<int>(nbricks);
    bp.list = dalloc<int>((size_t)nbricks * cap);
    bp.plane_major = plane_fft() ? 1 : 0;
  }

  const double4 *pme_positions() const { return pme_atoms ? d_xpa : d_xpq; }

  // Lists the atoms of every brick their stencils touch.
  template <int N, int TL>
  void bin_t() {
    pme_bin<N, TL><<<dim3(blocks_for(pme_atoms ? n : max_places), 1, nz), 256, 0, pstream>>>(
        pme_atoms ? d_pme_n : d_counts, pme_positions(), pmp, bp, d_flags, fs);
  }
  template <int N>
  void bin_n() {
    switch (brick_tl) {
      case 8: bin_t<N, 8>(); break;
      case 4: bin_t<N, 4>(); break;
      case 2: bin_t<N, 2>(); break;
      default: bin_t<N, 1>(); break;
    }
  }
  void pme_bin_launch() {
    if (pme_order == 4) bin_n<4>();
    else if (pme_order == 6) bin_n<6>();
    else bin_n<8>();
  }

  template <int N, int TL>
  void spread_rows_t() {
    constexpr int R = BrickShape<TL>::R, BX = BrickShape<TL>::BX, BY = BrickShape<TL>::BY;
    static_assert(BX * BY == R, "brick shape");
    const int blocks = ((pme_grid[0] + BX - 1) / BX) * ((pme_grid[1] + BY - 1) / BY);
    const size_t lds = (size_t)(32 * TL + 8) * pme_grid[2];
    pme_spread_rows<N, TL><<<dim3(blocks, 1, nz), 256, lds, pstream>>>(
        pme_positions(), pmp, bp, d_cplx, fft.ax[2].axis, fs);
  }
  template <int N>
  void spread_rows_n() {
    switch (brick_tl) {
      case 8: spread_rows_t<N, 8>(); break;
      case 4: spread_rows_t<N, 4>(); break;
      case 2: spread_rows_t<N, 2>(); break;
      default: spread_rows_t<N, 1>(); break;
    }
  }
  template <int N, int TL>
  void gather_rows_t() {
    constexpr int BX = BrickShape<TL>::BX, BY = BrickShape<TL>::BY;
    const int blocks = ((pme_grid[0] + BX - 1) / BX) * ((pme_grid[1] + BY - 1) / BY);
    const size_t lds = (size_t)(16 * TL + 8) * pme_grid[2];
    pme_gather_rows<N, TL><<<dim3(blocks, 1, nz), 256, lds, pstream>>>(
        pme_positions(), pmp, bp, d_cplx, fft.ax[2].axis, pme_atoms ? d_Fr : d_F, fs);
  }
  template <int N>
  void gather_rows_n() {
    switch (brick_tl) {
      case 8: gather_rows_t<N, 8>(); break;
      case 4: gather_rows_t<N, 4>(); break;
      case 2: gather_rows_t<N, 2>(); break;
      default: gather_rows_t<N, 1>(); break;
    }
  }
  // Backward z transform and gather, by the spread's bricks.
  void pme_gather_rows_launch() {
    if (pme_order == 4) gather_rows_n<4>();
    else if (pme_order == 6) gather_rows_n<6>();
    else gather_rows_n<8>();
  }

  // Spread and forward z transform: the half spectra rows of d_cplx.
  void pme_spread_launch() {
    if (pme_order == 4) spread_rows_n<4>();
    else if (pme_order == 6) spread_rows_n<6>();
    else spread_rows_n<8>();
  }


  // The y and x passes with the influence product: one kernel per kz plane
  // when a plane fits in LDS, else the y pass, the fused x pass and the y
  // pass again over tiles of lines.
  bool plane_fits() const {
    return use_pme && pme_grid[0] * pme_grid[1] <= kPlaneMax &&
           pme_grid[0] <= kPlaneAxisMax && pme_grid[1] <= kPlaneAxisMax;
  }
  // The plane kernels run when a plane fits, except the MFMA products of
  // edges other than 32 and 64, which take the tiles of lines (and with
  // them the [x][y][kz] half spectrum the spread and gather then write).
  bool plane_fft() const { return plane_fits() && !(pme_transform == 1 && !mfma_plane_edges()); }
  static constexpr int kPlaneMax = 4096;
  int recip_blocks() const {
    if (plane_fft()) return fft.H();
    return mfma_axis() ? pme_grid[1] * ((fft.H() + 31) / 32) : fft.x_blocks();
  }
  static bool mfma_edge(int k) { return k == 32 || k == 64; }
  bool mfma_plane_edges() const { return mfma_edge(pme_grid[0]) && mfma_edge(pme_grid[1]); }
  // The MFMA products take planes of edges 32 or 64 held in LDS, and any
  // other x and y edges up to kMfmaAxisMax over tiles of lines.
  bool mfma_possible() const {
    if (!use_pme) return false;
    if (plane_fits() && mfma_plane_edges()) return true;
    return pme_grid[0] <= kMfmaAxisMax && pme_grid[1] <= kMfmaAxisMax;
  }
  bool mfma_dft() const { return pme_transform == 1 && plane_fft(); }
  bool mfma_axis() const { return pme_transform == 1 && !plane_fft(); }
  static int pad32(int k) { return (k + 31) / 32 * 32; }

  // The DFT matrix W_K[n][k] = exp(-2 pi i n k / K) in three bf16 pieces
  // per part (h + m + l = W to 2^-24), laid out as fft_plane_conv_mfma reads
  // it: ((piece * 2 + part) * K + n) * K + k.
  static uint16_t bf16_bits(float v) {  // round to nearest even
    uint32_t u;
    memcpy(&u, &v, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
  }
  static float bf16_value(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float v;
    memcpy(&v, &u, 4);
    return v;
  }
  // Kp >= K: the table padded to Kp x Kp with zeros beyond K (the layout
  // of dft_axis_mfma; Kp = K for the plane kernel).
  static __bf16 *dft_table(int K, int Kp) {
    std::vector<uint16_t> t((size_t)6 * Kp * Kp, 0);
    for (int n = 0; n < K; ++n)
      for (int k = 0; k < K; ++k) {
        const long nk = ((long)n * k) % K;
        const double ang = -2.0 * M_PI * (double)nk / (double)K;
        const float part[2] = {(float)cos(ang), (float)sin(ang)};
        for (int c = 0; c < 2; ++c) {
          float r = part[c];
          for (int p = 0; p < 3; ++p) {
            const uint16_t b = bf16_bits(r);
            t[((size_t)(p * 2 + c) * Kp + n) * Kp + k] = b;
            r -= bf16_value(b);
          }
        }
      }
    __bf16 *d = dalloc<__bf16>(t.size());
    HIPCHK(hipMemcpy(d, t.data(), t.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    return d;
  }

  void set_pme_transform(int mode) {
    if (mode != 0 && mode != 1) throw HipError("PME transform: 0 (FFT) or 1 (MFMA DFT)");
    if (mode == 1 && !mfma_possible())
      throw HipError("the MFMA DFT takes PME grids with x and y edges of at most 224");
    if (mode == 1 && nz > 1 && !mfma_plane_edges())
      throw HipError("the MFMA DFT of frame batches takes x and y edges of 32 or 64");
    if (mode == 1 && !d_dft1) {
      const bool pad = !(plane_fits() && mfma_plane_edges());
      const int k0 = pme_grid[0], k1 = pme_grid[1];
      d_dft1 = dft_table(k0, pad ? pad32(k0) : k0);
      d_dft2 = k1 == k0 ? nullptr : dft_table(k1, pad ? pad32(k1) : k1);
    }
    const bool layout = plane_fft();
    pme_transform = mode;
    if (use_pme && plane_fft() != layout) make_bricks();  // the spread's half-spectrum layout
    invalidate_graph();
  }

  template <int K1, int K2>
  void plane_conv_mfma_t(cpx<float> *data, double *ep, double *vp) {
    const size_t lds = plane_mfma_lds(K1, K2);
    static bool attr = false;
    if (!attr) {
      HIPCHK(hipFuncSetAttribute((const void *)fft_plane_conv_mfma<K1, K2>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      attr = true;
    }
    const DftTables dt{d_dft1, d_dft2 ? d_dft2 : d_dft1};
    fft_plane_conv_mfma<K1, K2><<<dim3(fft.H(), 1, nz), 512, lds, pstream>>>(data, dt, ct, ep,
                                                                              vp, fs);
  }

  // y and x transforms with the influence product of every kz plane of
  // `data` (energy and virial partials when not null), by the chosen
  // transform.
  void plane_conv(cpx<float> *data, double *ep, double *vp) {
    if (mfma_dft()) {
      const int a = pme_grid[0], b = pme_grid[1];
      if (a == 64 && b == 64) plane_conv_mfma_t<64, 64>(data, ep, vp);
      else if (a == 32 && b == 32) plane_conv_mfma_t<32, 32>(data, ep, vp);
      else if (a == 64) plane_conv_mfma_t<64, 32>(data, ep, vp);
      else plane_conv_mfma_t<32, 64>(data, ep, vp);
      return;
    }
    const size_t lds = (size_t)2 * pme_grid[0] * pme_grid[1] * sizeof(cpx<float>);
    const int k0 = pme_grid[0], k1 = pme_grid[1];
    if (k0 == 64 && k1 == 64)
      fft_plane_conv<64, 64><<<dim3(fft.H(), 1, nz), 512, lds, pstream>>>(
          data, k0, k1, fft.ax[0].axis, fft.ax[1].axis, ct, ep, vp, fs);
    else if (k0 == 32 && k1 == 32)
      fft_plane_conv<32, 32><<<dim3(fft.H(), 1, nz), 512, lds, pstream>>>(
          data, k0, k1, fft.ax[0].axis, fft.ax[1].axis, ct, ep, vp, fs);
    else
      fft_plane_conv<0, 0><<<dim3(fft.H(), 1, nz), 512, lds, pstream>>>(
          data, k0, k1, fft.ax[0].axis, fft.ax[1].axis, ct, ep, vp, fs);
  }

  // The y transforms (dft sign +1 forward, -1 backward) and the x pass of
  // the convolution as MFMA DFT products, on grids whose planes do not fit.
  void axis_mfma(int axis, float sign, double *ep) {
    const int K = pme_grid[axis == 1 ? 1 : 0], Kp = pad32(K);
    const size_t lds = axis_mfma_lds(Kp);
    static bool attr = false;
    if (!attr) {
      HIPCHK(hipFuncSetAttribute((const void *)dft_axis_mfma<1>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)axis_mfma_lds(kMfmaAxisMax)));
      HIPCHK(hipFuncSetAttribute((const void *)dft_axis_mfma<0>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)axis_mfma_lds(kMfmaAxisMax)));
      attr = true;
    }
    const __bf16 *tab = axis == 1 && d_dft2 ? d_dft2 : d_dft1;
    const int tiles = (fft.H() + 31) / 32;
    const int lines = axis == 1 ? pme_grid[0] : pme_grid[1];
    if (axis == 1)
      dft_axis_mfma<1><<<dim3(lines * tiles, 1, nz), 512, lds, pstream>>>(
          d_cplx, pme_grid[0], pme_grid[1], fft.H(), Kp, tab, ct, sign, nullptr, fs);
    else
      dft_axis_mfma<0><<<dim3(lines * tiles, 1, nz), 512, lds, pstream>>>(
          d_cplx, pme_grid[0], pme_grid[1], fft.H(), Kp, tab, ct, 1.0f, ep, fs);
  }

  void pme_yx(bool energy) {
    double *ep = energy ? d_erecip : nullptr;
    if (plane_fft()) {
      plane_conv(d_cplx, ep, energy ? d_vrecip : nullptr);
      return;
    }
    if (mfma_axis()) {
      axis_mfma(1, 1.0f, nullptr);
      axis_mfma(0, 1.0f, ep);
      axis_mfma(1, -1.0f, nullptr);
      return;
    }
    const int TL = Fft3D<float>::TL;
    const int tiles = (fft.H() + TL - 1) / TL;
    fft_cols_y<float><<<fft.K[0] * tiles, 256, fft.lds(fft.K[1]), pstream>>>(
        d_cplx, fft.K[1], fft.H(), fft.ax[1].axis, -1);
    fft.x_pass(d_cplx, 0, ct, ep, pstream);
    fft_cols_y<float><<<fft.K[0] * tiles, 256, fft.lds(fft.K[1]), pstream>>>(
        d_cplx, fft.K[1], fft.H(), fft.ax[1].axis, +1);
  }

  void pme(bool energy) {
    if (!use_pme) return;
    pme_bin_launch();
    pme_spread_launch();
    pme_yx(energy);
    pme_gather_rows_launch();
  }

  // Pair loop on the engine's stream, PME on the side stream, joined.
  // With with_build, the list build (a no-op unless a rebuild was asked
  // for) comes first on the engine's stream: before the fork when PME reads
  // the place order the build writes, after it when PME runs in atom order
  // (pme_atoms), beside the build.  With timed >= 0, the pair kernel
  // records its span in slot `timed` of the timing buffer.
  void forces(bool energy, int timed = -1, bool with_build = false) {
    const bool recip = use_pme && (terms & 4);
    const bool split = recip && pme_atoms;
    if (with_build && !split) build();
    if (split) {
      // the list branch on the side stream, PME on the engine's stream
      // (launched first: the graph keeps the branch launched first on the
      // queue of the integration, which starts it sooner)
      HIPCHK(hipEventRecord(ev_fork, stream));
      HIPCHK(hipStreamWaitEvent(side, ev_fork, 0));
      const hipStream_t ps = pstream;
      pstream = stream;
      pme(energy);
      pstream = ps;
      lstream = side;
      if (with_build) build();
    } else if (recip) {
      HIPCHK(hipEventRecord(ev_fork, stream));
      HIPCHK(hipStreamWaitEvent(side, ev_fork, 0));
      pme(energy);
      HIPCHK(hipEventRecord(ev_join, side));
    }
    if (timed >= 0) stamp_clock<<<1, 64, 0, lstream>>>(d_timing + 3 * timed);
    pairs(energy);
    if (timed >= 0) {
      stamp_clock<<<1, 64, 0, lstream>>>(d_timing + 3 * timed + 1);
      stamp_clock<<<1, 64, 0, lstream>>>(d_timing + 3 * timed + 2);
    }
    if (split) {
      HIPCHK(hipEventRecord(ev_join, side));
      lstream = stream;
    }
    if (recip) HIPCHK(hipStreamWaitEvent(stream, ev_join, 0));
  }

  void clear_forces() {
    clear_words<<<dim3(256, 1, nz), 256, 0, stream>>>(d_F, 3 * (size_t)max_places, fs);
    if (d_Fr) clear_words<<<dim3(64, 1, nz), 256, 0, stream>>>(d_Fr, 3 * (size_t)n, 0);
  }

  // Outside the step sequence (evaluation, minimisation): PME in atom order
  // reads the positions the integrator would have written, and its words
  // are folded into d_F afterwards, so that every reader of d_F sees the
  // whole force.
  void pack_pme_atoms() {
    if (d_xpa) pme_pack_atoms<<<blocks_for(n), 256, 0, stream>>>(n, d_x, d_q, d_xpa);
  }
  void fold_pme_atoms() {
    if (d_Fr)
      pme_fold_atoms<<<blocks_for(n), 256, 0, stream>>>(n, d_place_of, d_Fr, d_F);
  }

  template <int PART>
  void integrate(double dt) {
    ip.dt = dt;
    ip.langevin = lv_friction > 0.0;
    if (ip.langevin) {
      const double kB = 0x1.1072c483af26dp-7;  // kJ mol^-1 K^-1 (SI k_B N_A)
      ip.decay = exp(-lv_friction * dt);
      ip.spread = sqrt((1.0 - ip.decay * ip.decay) * kB * lv_temperature);
      ip.seed = lv_seed;
      ip.step = d_lstep;
    } else {
      ip.step = nullptr;
    }
    if (nclusters)
      integrate_clusters<PART><<<blocks_for(nclusters, 64), 64, 0, stream>>>(
          nclusters, d_cl_first, d_cl_code, d_x, d_v, d_inv_mass, d_F, d_Fr, d_place_of,
          d_shift, d_xref, d_q, d_xq, use_origins() ? d_gorig : nullptr,
          d_xpa ? nullptr : d_xpq, d_xpa, ip, d_flags);
    // at least one block: its last block advances the Langevin step number
    integrate_part<PART><<<std::max(blocks_for(nunits, 64), 1), 64, 0, stream>>>(
        nunits, d_unit_first, d_unit_size, d_x, d_v, d_inv_mass, d_F, d_Fr, d_place_of,
        d_shift, d_xref, d_q, d_xq, use_origins() ? d_gorig : nullptr,
        d_xpa ? nullptr : d_xpq, d_xpa, ip, d_flags);
  }
  void step_first(double dt) { integrate<kFirstHalf>(dt); }
  void step_second(double dt) { integrate<kSecondHalf>(dt); }
  void step_between(double dt) { integrate<kBetween>(dt); }

  // `count` whole steps; between two steps the second half of one and the
  // first half of the next run as one kernel.
  void steps_seq(int64_t count, double dt) {
    if (count <= 0) return;
    step_first(dt);
    for (int64_t s = 0; s < count; ++s) {
      forces(false, measure ? (int)s : -1, true);
      if (s + 1 < count) step_between(dt);
      else step_second(dt);
    }
  }

  // Live timing of the pair loop inside the steps: device-clock readings
  // just before and after the pair kernel of step s land in slots 3 s and
  // 3 s + 1, and a third right after the second in 3 s + 2 (in the captured
  // graphs too): the pair kernel took (t1 - t0) - (t2 - t1), the second
  // difference being what one reading adds.  run() resets the slots before
  // each chunk and reads them after its flag check.
  bool measure = false;
  unsigned long long *d_timing = nullptr;
  int64_t timing_slots = 0;
  double clock_khz = 0.0;
  double pair_ms_sum = 0.0;
  int64_t pair_ms_count = 0;
  void ensure_timing(int64_t count) {
    if (timing_slots < count) {
      dfree(d_timing);
      d_timing = dalloc<unsigned long long>(3 * (size_t)count);
      timing_slots = count;
      invalidate_graph();  // captured graphs point at the old buffer
    }
    if (clock_khz == 0.0) {
      int dev = 0, khz = 0;
      HIPCHK(hipGetDevice(&dev));
      HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
      if (khz <= 0) throw HipError("the device reports no wall-clock rate");
      clock_khz = (double)khz;
    }
  }
  void reset_timing(int64_t count) {
    std::vector<unsigned long long> init(3 * (size_t)count, 0ull);
    to_device(d_timing, init.data(), init.size() * sizeof(unsigned long long));
  }
  void collect_pair_times(int64_t count) {
    std::vector<unsigned long long> t(3 * (size_t)count);
    to_host(t.data(), d_timing, t.size() * sizeof(unsigned long long));
    for (int64_t s = 0; s < count; ++s) {
      const unsigned long long t0 = t[3 * s], t1 = t[3 * s + 1], t2 = t[3 * s + 2];
      if (t0 == 0 || t1 < t0 || t2 < t1) continue;  // not launched
      pair_ms_sum += ((double)(t1 - t0) - (double)(t2 - t1)) / clock_khz;
      ++pair_ms_count;
    }
  }

  std::vector<int> read_flags() {
    std::vector<int> f(kNumFlags);
    HIPCHK(hipMemcpyAsync(f.data(), d_flags, kNumFlags * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return f;
  }

  void clear_error_flags() {
    std::vector<int> z(kNumFlags, 0);
    HIPCHK(hipMemcpyAsync(d_flags, z.data(), kNumFlags * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }

  // Overflow of a column bucket or a row: grow and report true.
  bool grow_if_overflowed(const std::vector<int> &f) {
    if (f[kFlagBrickOverflow]) {  // a brick listed more atoms than it holds
      make_bricks(bp.cap * 2);
      clear_error_flags();
      return true;
    }
    if (f[kFlagOverflowRow] & 2) throw HipError("more than 16 exclusions of one particle");
    if (!f[kFlagOverflowColumn] && !f[kFlagOverflowRow]) return false;
    int col_min = f[kFlagOverflowColumn] ? g.col_cap * 2 : g.col_cap;
    int row_min = (f[kFlagOverflowRow] & 1) ? g.row_cap * 2 : g.row_cap;
    int ex_min = (f[kFlagOverflowRow] & 4) ? ex_cap * 2 : ex_cap;
    if (col_min > kSortCapMax && f[kFlagOverflowColumn])
      throw HipError("a column holds more than 1024 particles");
    configure_list(list_reach, col_min, row_min, ex_min);
    clear_error_flags();
    return true;
  }

  void evaluate(bool energy) {
    for (int attempt = 0; attempt < 8; ++attempt) {
      clear_forces();
      request_build();
      pack_pme_atoms();
      forces(energy, -1, true);
      fold_pme_atoms();
      const std::vector<int> f = read_flags();
      if (!grow_if_overflowed(f)) {
        forces_valid = true;
        return;
      }
    }
    throw HipError("the neighbor list kept overflowing");
  }

  // Forces and energies of the terms `mask` names (the others are 0).
  void compute_terms(int mask, double *F_out, double *E_out) {
    if (mask < 1 || mask > 15) throw HipError("terms: a nonzero mask of MDRT_TERM_* bits");
    terms = mask;
    try {
      compute(F_out, E_out);
    } catch (...) {
      terms = 15;
      throw;
    }
    terms = 15;
    forces_valid = false;
  }

  void compute(double *F_out, double *E_out) {
    if (list_reach != cutoff + skin) configure_list(cutoff + skin);
    evaluate(true);
    outputs(F_out, E_out);
  }

  // Forces (atom order, f64) and energy terms of the last evaluation.
  void outputs(double *F_out, double *E_out) {
    if (F_out) {
      double *d_out = dalloc<double>(3 * (size_t)n);
      forces_to_atoms<<<blocks_for(n), 256, 0, stream>>>(n, d_place_of, d_F, d_out, 0);
      HIPCHK(hipMemcpyAsync(F_out, d_out, 3 * n * sizeof(double), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      dfree(d_out);
    }
    if (E_out) {
      for (int k = 0; k < MDRT_NUM_ENERGIES; ++k) E_out[k] = 0.0;
      // one partial per (segment, group) wave, summed in a fixed order
      std::vector<double> ep(2 * (size_t)max_groups * nseg);
      to_host(ep.data(), d_epair, ep.size() * sizeof(double));
      for (size_t w = 0; (terms & 1) && w < ep.size() / 2; ++w) {
        E_out[MDRT_E_LJ] += ep[2 * w];
        E_out[MDRT_E_COUL_DIRECT] += ep[2 * w + 1];
      }
      if (nspecial && (terms & 2)) {
        const int nb = blocks_for(nspecial);
        std::vector<double> es(3 * (size_t)nb);
        to_host(es.data(), d_espec, es.size() * sizeof(double));
        for (int b = 0; b < nb; ++b) {
          E_out[MDRT_E_COUL_EXCL] += es[3 * b];
          E_out[MDRT_E_LJ14] += es[3 * b + 1];
          E_out[MDRT_E_COUL14] += es[3 * b + 2];
        }
      }
      if (const int tb = tuple_blocks()) {
        std::vector<double> et((size_t)kTupleKinds * tb);
        to_host(et.data(), d_etup, et.size() * sizeof(double));
        const int slot[kTupleKinds] = {MDRT_E_BONDS, MDRT_E_UREY_BRADLEY, MDRT_E_ANGLES,
                                       MDRT_E_DIHEDRALS, MDRT_E_IMPROPERS};
        for (int b = 0; b < tb; ++b)
          for (int k = 0; k < kTupleKinds; ++k) E_out[slot[k]] += et[(size_t)kTupleKinds * b + k];
      }
      if (use_pme && (terms & 4)) {
        std::vector<double> er(recip_blocks());
        to_host(er.data(), d_erecip, er.size() * sizeof(double));
        for (double e : er) E_out[MDRT_E_COUL_RECIP] += e;
      }
      E_out[MDRT_E_COUL_SELF] = e_self;
      E_out[MDRT_E_DISP_CORR] = e_disp;
      E_out[MDRT_E_KINETIC] = kinetic_energy();
    }
  }

  // The unordered atom pairs (i < j) of `rows` whose f32 minimum-image
  // distance at the current positions is within `far`, as sorted unique
  // keys (i << 32 | j); explicit masks drop the excluded pairs.
  std::vector<long long> export_pairs(const PairEntry *rows, const int *row_count, double far,
                                      int64_t capacity_hint) {
    const float limit2 = (float)(far * far);
    const unsigned long long cap2 =
        (unsigned long long)std::max<int64_t>(capacity_hint, 1) * 4 + 1024;
    int2 *d_pairs = dalloc<int2>(cap2);
    unsigned long long *d_count = dalloc<unsigned long long>(1);
    nb_export_pairs<<<blocks_for(max_groups, 4), 256, 0, stream>>>(
        d_x, d_atom_at, d_real_mask, rows, row_count, d_ex, d_counts, g.row_cap, (float)box[0],
        (float)box[1], (float)box[2], box[0], box[1], box[2], limit2, d_pairs, d_count, cap2);
    unsigned long long cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, d_count, sizeof(cnt), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (cnt > cap2) {
      dfree(d_pairs);
      dfree(d_count);
      return export_pairs(rows, row_count, far, (int64_t)cnt);
    }
    std::vector<int2> h(cnt);
    if (cnt) to_host(h.data(), d_pairs, cnt * sizeof(int2));
    dfree(d_pairs);
    dfree(d_count);
    std::vector<long long> keys(cnt);
    for (size_t i = 0; i < cnt; ++i) keys[i] = ((long long)h[i].x << 32) | (unsigned)h[i].y;
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    return keys;
  }

  // The check of the list between rebuilds (the counterpart of the
  // reference's MDIR_CHECK_PRUNING, runtime/mdrt.c:48-79): the forces the
  // next step would compute at the current positions, through the list the
  // steps hold (the build kernels run only if the last drift asked for a
  // rebuild, as in a step), and the pairs within rc found in that list and
  // in the rows the pair loop runs over (the pruned rows with a dual list).
  // counts[0], counts[1]: those pairs; counts[2]: list builds the check
  // itself did (0 unless the last drift asked for one).
  void check_list(double *F_out, double *E_out, int64_t *counts) {
    if (!forces_valid || list_reach != cutoff + skin)
      throw HipError("no list of a run to check: the positions changed since the last step");
    const int builds0 = read_flags()[kFlagBuilds];
    clear_forces();
    pack_pme_atoms();
    forces(true, -1, true);
    fold_pme_atoms();
    const std::vector<int> f = read_flags();
    if (grow_if_overflowed(f)) {
      forces_valid = false;
      throw HipError("the list overflowed at the check");
    }
    outputs(F_out, E_out);
    if (counts) {
      const double far = cutoff + 3.0e-6 * ((box[0] + box[1]) + box[2]);
      const int64_t hint = 8 * (int64_t)n;
      counts[0] = (int64_t)export_pairs(d_rows, d_row_count, far, hint).size();
      counts[1] = prune_skin > 0
                      ? (int64_t)export_pairs(d_rows_in, d_row_count_in, far, hint).size()
                      : counts[0];
      counts[2] = f[kFlagBuilds] - builds0;
    }
  }

  double kinetic_energy() {
    std::vector<double> v(3 * (size_t)n);
    to_host(v.data(), d_v, v.size() * sizeof(double));
    double ke = 0;
    for (int i = 0; i < n; ++i)
      ke += 0.5 * h_mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    return ke;
  }

  // The captured sequence of `steps` steps of dt (with the pair-loop clock
  // readings when `measure` is set), from the cache or captured now.
  hipGraphExec_t capture(int steps, double dt) {
    if (hipGraphExec_t e = find_graph(steps, dt, measure)) return e;
    if (step_graphs.size() >= kMaxStepGraphs) {
      hipGraphExecDestroy(step_graphs.front().exec);
      step_graphs.erase(step_graphs.begin());
    }
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    steps_seq(steps, dt);
    HIPCHK(hipStreamEndCapture(stream, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t st = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    HIPCHK(hipGraphDestroy(graph));
    HIPCHK(st);
    step_graphs.push_back({steps, dt, measure, exec});
    return exec;
  }

  // Captures, without running them, the graphs a run of `steps` steps in
  // chunks of `graph` replays: the graph of `graph` steps and that of the
  // remainder.  Capture time then stays out of a timed run.
  void prepare(int64_t steps, double dt, int graph) {
    if (graph <= 0 || steps <= 0) return;
    if (list_reach != cutoff + skin) configure_list(cutoff + skin);
    if (!forces_valid) evaluate(false);
    if (steps >= graph) capture(graph, dt);
    if (steps % graph) capture((int)(steps % graph), dt);
    HIPCHK(hipStreamSynchronize(stream));
  }

  // Steps in chunks (one graph launch, or up to 64 single steps without
  // graphs); after each chunk the flags are read, and a chunk whose list
  // overflowed is redone from its saved state with the grown list.
  void run(int64_t steps, double dt, int graph, double *kinetic, double *pair_ms = nullptr) {
    if (list_reach != cutoff + skin) configure_list(cutoff + skin);
    if (!forces_valid) evaluate(false);
    const int64_t chunk = graph > 0 ? graph : 64;
    measure = pair_ms != nullptr;
    pair_ms_sum = 0.0;
    pair_ms_count = 0;
    if (measure) ensure_timing(chunk);
    int64_t done = 0;
    while (done < steps) {
      const int64_t todo = std::min(chunk, steps - done);
      for (int attempt = 0;; ++attempt) {
        if (measure) reset_timing(todo);
        put_step_count();
        on_device(d_save_x, d_x, 3 * n * sizeof(double));
        on_device(d_save_v, d_v, 3 * n * sizeof(double));
        if (graph > 0) {
          HIPCHK(hipGraphLaunch(capture((int)todo, dt), stream));
        } else {
          steps_seq(todo, dt);
        }
        const std::vector<int> f = read_flags();  // synchronises the stream
        if (!grow_if_overflowed(f)) {
          if (measure) collect_pair_times(todo);
          break;
        }
        if (attempt == 7) throw HipError("the neighbor list kept overflowing");
        on_device(d_x, d_save_x, 3 * n * sizeof(double));
        on_device(d_v, d_save_v, 3 * n * sizeof(double));
        evaluate(false);
      }
      done += todo;
      step_count += todo;
    }
    HIPCHK(hipStreamSynchronize(stream));
    measure = false;
    if (pair_ms) *pair_ms = pair_ms_count ? pair_ms_sum / (double)pair_ms_count : 0.0;
    if (kinetic) *kinetic = kinetic_energy();
  }

  // ---- launch-shape trial ----
  // Times the launch shapes of this system's kernels and keeps the fastest
  // (the counterpart of the reference's launch-shape trial,
  // tools/mdir/TowersTrial.h:13-33, Run.cpp:913-926). Candidates: with
  // `what` & 1, PME bricks of 8, 4, 2 or 1 complex z lines (2 TL grid rows;
  // those whose LDS fits); with & 2, the plane transform (Stockham passes or
  // the split-BF16x3 MFMA products, where the grid allows them); with & 4,
  // pair-loop segments of 16, 24 or 48 entries. Each candidate runs from
  // the same state: one untimed graph of `steps` steps (list build,
  // capture), then `graphs` graphs timed on the host clock around run(),
  // overflow checks included; two passes over the candidates, the faster of
  // each candidate's two times counting. The current shape is kept unless
  // another is at least kTrialMargin faster. Positions and velocities are
  // restored at the end. ms_out[(8 * s + 4 * transform) + k] = ms per step
  // with segments of kSegLens[s] and bricks of TL = 8 >> k (NaN: not run).
  // The shapes give the same forces up to the f32 rounding of the gather's
  // per-brick partial sums (up to about 1e-7 relative) and the order of the
  // pair loop's per-wave energy partials; a fixed choice is bit-exact run
  // to run.
  static constexpr int kTrialShapes = 24;
  static constexpr int kSegLens[3] = {16, 24, 48};
  static constexpr double kTrialMargin = 0.02;
  struct Shape {
    int s, tr, k;  // segment index, transform, brick index (TL = 8 >> k)
    int slot() const { return 8 * s + 4 * tr + k; }
  };
  static int seg_index(int len) { return len == 16 ? 0 : (len == 24 ? 1 : 2); }
  static std::vector<Shape> trial_shapes(int what, int K2, bool mfma_ok, int tl0, int tr0,
                                         int seg0) {
    std::vector<Shape> c;
    for (int si = 0; si < 3; ++si) {
      if (!(what & 4) && kSegLens[si] != seg0) continue;
      for (int tr = 0; tr < 2; ++tr) {
        if (!(what & 2) && tr != tr0) continue;
        if (tr == 1 && !mfma_ok) continue;
        for (int k = 0; k < 4; ++k) {
          const int tl = 8 >> k;
          if (!(what & 1) && tl != tl0) continue;
          if (K2 > 0 && (size_t)(32 * tl + 8) * K2 > 60 * 1024) continue;
          c.push_back({si, tr, k});
        }
      }
    }
    return c;
  }
  // Runs measure(shape) for every candidate in two passes, fills ms_out
  // with the faster time of each, and returns the shape to keep.
  template <class Measure>
  static Shape trial(const std::vector<Shape> &cands, Shape cur, Measure measure,
                     double *ms_out) {
    for (int pass = 0; pass < 2; ++pass)
      for (const Shape &c : cands) {
        const double ms = measure(c);
        double &slot = ms_out[c.slot()];
        slot = std::isnan(slot) ? ms : std::min(slot, ms);
      }
    Shape best = cur;
    bool have = false, cur_in = false;
    for (const Shape &c : cands) {
      if (!have || ms_out[c.slot()] < ms_out[best.slot()]) best = c;
      have = true;
      if (c.slot() == cur.slot()) cur_in = true;
    }
    if (cur_in && ms_out[best.slot()] > (1.0 - kTrialMargin) * ms_out[cur.slot()]) best = cur;
    return best;
  }
  // A segment length: the list layout follows (reallocated, rebuilt at the
  // next evaluation).
  void set_segment(int len) {
    if (len != 16 && len != 24 && len != 48)
      throw HipError("pair-loop segments: 16, 24 or 48 entries");
    seg_len = len;
    if (d_rows && seg_len_cfg != len) configure_list(list_reach, g.col_cap, g.row_cap, ex_cap);
  }
  // One launch shape, set directly (what the trial keeps can be set again
  // without timing): segment entries 16, 24 or 48; PME bricks of 8, 4, 2 or
  // 1 z lines that fit LDS; plane transform 0 or 1 where the grid allows.
  void set_launch_shape(int seg, int tl, int tr) {
    if (tl != 8 && tl != 4 && tl != 2 && tl != 1) throw HipError("PME bricks: 8, 4, 2 or 1 z lines");
    if (use_pme && (size_t)(32 * tl + 8) * pme_grid[2] > 60 * 1024)
      throw HipError("PME bricks of that many z lines exceed LDS at this grid");
    set_segment(seg);
    if (use_pme) {
      brick_tl_max = tl;
      make_bricks();
      set_pme_transform(tr);
    } else if (tr != 0) {
      throw HipError("no PME: the plane transform is 0");
    }
    forces_valid = false;
    invalidate_graph();
  }
  void get_launch_shape(int32_t *out) const {
    out[0] = seg_len;
    out[1] = use_pme ? brick_tl : 0;
    out[2] = pme_transform;
  }

  void autotune(int what, int steps, int graphs, double dt, double *ms_out) {
    for (int k = 0; k < kTrialShapes; ++k) ms_out[k] = NAN;
    if (what <= 0 || steps <= 0 || graphs <= 0) return;
    const size_t bytes = 3 * (size_t)n * sizeof(double);
    std::vector<double> x0(3 * (size_t)n), v0(3 * (size_t)n);
    to_host(x0.data(), d_x, bytes);
    to_host(v0.data(), d_v, bytes);
    const int64_t step0 = step_count;
    auto restore = [&]() {
      step_count = step0;
      HIPCHK(hipMemcpyAsync(d_x, x0.data(), bytes, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(d_v, v0.data(), bytes, hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
      forces_valid = false;
    };
    if (!use_pme) what &= 4;  // no PME shapes to try
    const bool mfma_ok = mfma_possible();
    const int tl0 = use_pme ? brick_tl : 8, tr0 = pme_transform;
    const Shape cur{seg_index(seg_len), tr0, tl0 == 8 ? 0 : (tl0 == 4 ? 1 : (tl0 == 2 ? 2 : 3))};
    auto apply = [&](const Shape &c) {
      set_segment(kSegLens[c.s]);
      if (use_pme) {
        brick_tl_max = 8 >> c.k;
        make_bricks();
        set_pme_transform(c.tr);
      }
    };
    const Shape keep = trial(
        trial_shapes(what, use_pme ? pme_grid[2] : 0, mfma_ok, tl0, tr0, seg_len), cur,
        [&](const Shape &c) {
          apply(c);
          restore();
          run(steps, dt, steps, nullptr);
          const auto t0 = std::chrono::steady_clock::now();
          run((int64_t)steps * graphs, dt, steps, nullptr);
          const auto t1 = std::chrono::steady_clock::now();
          return std::chrono::duration<double, std::milli>(t1 - t0).count() /
                 ((double)steps * graphs);
        },
        ms_out);
    apply(keep);
    restore();
  }

  void minimize(int64_t steps, double max_step, double *energy_out) {
    double E[MDRT_NUM_ENERGIES];
    std::vector<double> F(3 * (size_t)n);
    compute(F.data(), E);
    auto potential = [](const double *e) { return potential_of(e); };
    double e0 = potential(E);
    double h = max_step;
    double *d_x0 = dalloc<double>(3 * (size_t)n);
    for (int64_t s = 0; s < steps; ++s) {
      double fmax = 0;
      for (double f : F) fmax = std::max(fmax, fabs(f));
      if (fmax == 0) break;
      HIPCHK(hipMemcpyAsync(d_x0, d_x, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, stream));
      // the force words of the current positions are still in d_F
      descend<<<blocks_for(nunits), 256, 0, stream>>>(nunits, d_unit_first, d_unit_size, d_x0,
                                                      d_x, d_F, d_place_of, h / fmax, ip);
      if (nclusters)
        descend<<<blocks_for(nclusters), 256, 0, stream>>>(nclusters, d_cl_first, d_cl_code,
                                                           d_x0, d_x, d_F, d_place_of, h / fmax,
                                                           ip);
      std::vector<double> F1(3 * (size_t)n);
      double E1[MDRT_NUM_ENERGIES];
      double e1;
      try {
        compute(F1.data(), E1);
        e1 = potential(E1);
      } catch (const HipError &) {
        // a trial too wild to evaluate (a list that cannot hold it) is a
        // rejected step
        clear_error_flags();
        e1 = INFINITY;
      }
      if (e1 < e0) {
        e0 = e1;
        F.swap(F1);
        h = std::min(1.2 * h, max_step);  // steepest descent: grow, capped
      } else {
        HIPCHK(hipMemcpyAsync(d_x, d_x0, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, stream));
        compute(F.data(), E);
        h *= 0.2;
      }
    }
    dfree(d_x0);
    forces_valid = false;
    if (energy_out) *energy_out = e0;
  }

  int64_t neighbor_pairs(double reach, int excluded, int64_t capacity, int32_t *out) {
    const double margin = 3.0e-6 * ((box[0] + box[1]) + box[2]);
    for (int k = 0; k < 3; ++k)
      if (box[k] <= 2.0 * (reach + margin + 1e-4))
        throw HipError("the reach must be below half of every box edge");
    // a list without exclusions when they are kept
    int *keep_start = d_excl_start;
    int *empty = nullptr;
    if (!excluded) {
      empty = dalloc<int>(n + 1);
      d_excl_start = empty;
    }
    configure_list(reach + margin);
    std::vector<int> f;
    for (int attempt = 0; attempt < 8; ++attempt) {
      request_build();
      build();
      clear_forces();
      pairs(false);  // clears the request
      f = read_flags();
      if (!grow_if_overflowed(f)) break;
    }
    d_excl_start = keep_start;
    dfree(empty);
    const std::vector<long long> keys =
        export_pairs(d_rows, d_row_count, reach + margin, capacity);
    const int64_t total = (int64_t)keys.size();
    for (int64_t i = 0; i < std::min(total, capacity); ++i) {
      out[2 * i] = (int32_t)(keys[i] >> 32);
      out[2 * i + 1] = (int32_t)(keys[i] & 0xffffffff);
    }
    configure_list(cutoff + skin);
    forces_valid = false;
    return total;
  }

  void stats(int64_t *s) {
    std::vector<int> counts(4);
    to_host(counts.data(), d_counts, 4 * sizeof(int));
    std::vector<int> rc(max_groups);
    to_host(rc.data(), d_row_count, rc.size() * sizeof(int));
    std::vector<PairEntry> rows((size_t)max_groups * g.row_cap);
    to_host(rows.data(), d_rows, rows.size() * sizeof(PairEntry));
    int64_t tiles = 0, largest = 0;
    for (int gi = 0; gi < counts[1] / kGroup; ++gi) {
      for (int e = 0; e < rc[gi]; ++e)
        tiles += __builtin_popcount((rows[(size_t)gi * g.row_cap + e].info >> 8) & ((1u << kGroup) - 1u));
      largest = std::max<int64_t>(largest, rc[gi]);
    }
    s[0] = counts[0];
    s[1] = counts[1];
    s[2] = tiles;
    s[3] = largest;
    s[4] = g.row_cap;
    s[5] = g.ncol;
    s[6] = g.col_cap;
    std::vector<int> f = read_flags();
    s[7] = f[kFlagBuilds];
    s[8] = tiles;
    s[9] = f[kFlagPrunes];
    if (prune_skin > 0) {
      std::vector<int> rci(max_groups);
      to_host(rci.data(), d_row_count_in, rci.size() * sizeof(int));
      std::vector<PairEntry> rin((size_t)max_groups * g.row_cap);
      to_host(rin.data(), d_rows_in, rin.size() * sizeof(PairEntry));
      int64_t t = 0;
      for (int gi = 0; gi < counts[1] / kGroup; ++gi)
        for (int e = 0; e < rci[gi]; ++e)
          t += __builtin_popcount((rin[(size_t)gi * g.row_cap + e].info >> 8) & ((1u << kGroup) - 1u));
      s[8] = t;
    }
  }

  // Device-event timing of every kernel of one step, on the stream the
  // kernels run on: ms per launch, averaged over `repeats` steps.
  //  0 integrate_first  1 nb_bin  2 nb_sort  3 nb_search (+ nb_prune with a pruned list)
  //  4 nb_pairs (with the excluded and 1-4 pairs)  5 pme_bin
  //  6 pme_spread_rows (spread + forward z transform)
  //  7 pme_yx (y and x transforms and the influence product)
  //  8 pme_gather_rows (backward z transform + gather)  9 integrate_second
  static constexpr int kProfileSlots = 10;
  void profile(int repeats, double *ms) {
    if (!forces_valid) evaluate(false);
    pstream = stream;  // kernel by kernel on one stream
    hipEvent_t ev[kProfileSlots + 1];
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    double acc[kProfileSlots] = {0};
    for (int r = 0; r < repeats; ++r) {
      int k = 0;
      HIPCHK(hipEventRecord(ev[k++], stream));
      step_first(0.002);
      HIPCHK(hipEventRecord(ev[k++], stream));
      launch_bin();
      HIPCHK(hipEventRecord(ev[k++], stream));
      launch_sort();
      HIPCHK(hipEventRecord(ev[k++], stream));
      launch_search();
      launch_prune();  // the pruned rows the pair kernel reads (if any)
      HIPCHK(hipEventRecord(ev[k++], stream));
      pairs(false);
      HIPCHK(hipEventRecord(ev[k++], stream));
      if (use_pme) pme_bin_launch();
      HIPCHK(hipEventRecord(ev[k++], stream));
      if (use_pme) pme_spread_launch();
      HIPCHK(hipEventRecord(ev[k++], stream));
      if (use_pme) pme_yx(false);
      HIPCHK(hipEventRecord(ev[k++], stream));
      if (use_pme) pme_gather_rows_launch();
      HIPCHK(hipEventRecord(ev[k++], stream));
      step_second(0.002);
      HIPCHK(hipEventRecord(ev[k++], stream));
      HIPCHK(hipEventSynchronize(ev[kProfileSlots]));
      for (int j = 0; j < kProfileSlots; ++j) {
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, ev[j], ev[j + 1]));
        acc[j] += t;
      }
    }
    for (int j = 0; j < kProfileSlots; ++j) ms[j] = acc[j] / repeats;
    for (auto &e : ev) hipEventDestroy(e);
    pstream = side;
  }
};

}  // namespace mdhip

#include "frames.hip"

using mdhip::Engine;
using mdhip::FrameBatch;
using mdhip::g_last_error;

template <typename F>
static int guarded(F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return -1;
  }
}

// ---- FFT plans of the runtime ABI (mdrt_hip.c) -------------------------

namespace {
struct AbiPlan {
  int64_t k[3];
  int single;
  int users;
  mdhip::Fft3D<float> f32;
  mdhip::Fft3D<double> f64;
};
std::vector<AbiPlan *> g_plans;

AbiPlan *find_plan(int64_t k1, int64_t k2, int64_t k3, int single) {
  for (AbiPlan *p : g_plans)
    if (p->k[0] == k1 && p->k[1] == k2 && p->k[2] == k3 && p->single == single) return p;
  AbiPlan *p = new AbiPlan();
  p->k[0] = k1;
  p->k[1] = k2;
  p->k[2] = k3;
  p->single = single;
  p->users = 0;
  const int K[3] = {(int)k1, (int)k2, (int)k3};
  try {
    if (single)
      p->f32.make(K);
    else
      p->f64.make(K);
  } catch (...) {
    delete p;
    throw;
  }
  g_plans.push_back(p);
  return p;
}
}  // namespace

// The serial number of a staging end, stored by the device into the slot's
// mapped host word once the copies before it on the stream are done: a
// vector store with a system-scope release, so that the host sees the
// copies' bytes before it sees the number.
__global__ void stage_mark(unsigned int *word, unsigned int serial) {
  __hip_atomic_store(word, serial, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

extern "C" int mdhip_stage_mark(void *word, uint32_t serial, void *stream) {
  stage_mark<<<1, 1, 0, (hipStream_t)stream>>>(static_cast<unsigned int *>(word), serial);
  return (int)hipGetLastError();
}

extern "C" int mdhip_fft_execute(int direction, int single, void *in, void *out, int64_t k1,
                                 int64_t k2, int64_t k3, const float *tables, void *stream) {
  return guarded([&] {
    AbiPlan *p = find_plan(k1, k2, k3, single);
    hipStream_t s = (hipStream_t)stream;
    mdhip::ConvTables ct{};
    const int kmax = (int)std::max({k1, k2, k3});
    if (tables)
      for (int a = 0; a < 3; ++a) {
        ct.t[a] = tables + a * kmax;
        ct.m[a] = tables + (3 + a) * kmax;
      }
    ct.K3 = (int)k3;
    if (single) {
      auto &f = p->f32;
      if (direction < 0) {
        f.forward(in, false, (mdhip::cpx<float> *)out, s);
        f.x_pass((mdhip::cpx<float> *)out, -1, ct, nullptr, s);
      } else {
        f.x_pass((mdhip::cpx<float> *)in, tables ? 2 : 1, ct, nullptr, s);
        f.backward((mdhip::cpx<float> *)in, (float *)out, s);
      }
    } else {
      if (tables) throw mdhip::HipError("the Gaussian backward transform is f32 only");
      auto &f = p->f64;
      if (direction < 0) {
        f.forward(in, false, (mdhip::cpx<double> *)out, s);
        f.x_pass((mdhip::cpx<double> *)out, -1, ct, nullptr, s);
      } else {
        f.x_pass((mdhip::cpx<double> *)in, 1, ct, nullptr, s);
        f.backward((mdhip::cpx<double> *)in, (double *)out, s);
      }
    }
    HIPCHK(hipGetLastError());
  });
}

extern "C" int mdhip_fft_retain(int64_t k1, int64_t k2, int64_t k3, int single) {
  int handle = -1;
  if (guarded([&] {
        AbiPlan *p = find_plan(k1, k2, k3, single);
        ++p->users;
        for (size_t i = 0; i < g_plans.size(); ++i)
          if (g_plans[i] == p) handle = (int)i;
      }) != 0)
    return -1;
  return handle;
}

extern "C" void mdhip_fft_release(int handle) {
  if (handle >= 0 && handle < (int)g_plans.size() && g_plans[handle]->users > 0)
    --g_plans[handle]->users;
}

extern "C" int64_t mdhip_fft_plan_count(void) { return (int64_t)g_plans.size(); }

extern "C" int64_t mdhip_fft_plan_bytes(void) {
  int64_t b = 0;
  for (AbiPlan *p : g_plans)
    for (int a = 0; a < 3; ++a) b += p->k[a] * (p->single ? 8 : 16);
  return b;
}

extern "C" void mdhip_fft_trim(void) {
  std::vector<AbiPlan *> keep;
  for (AbiPlan *p : g_plans) {
    if (p->users) {
      keep.push_back(p);
      continue;
    }
    p->f32.release();
    p->f64.release();
    delete p;
  }
  g_plans.swap(keep);
}


extern "C" {

const char *mdrtHipLastError(void) { return g_last_error.c_str(); }

void *mdrtHipSystemCreate(const mdrtHipSystemDesc *desc) {
  Engine *e = new Engine();
  if (guarded([&] { e->create(desc); }) != 0) {
    delete e;
    return nullptr;
  }
  return e;
}

void mdrtHipSystemDestroy(void *system) { delete static_cast<Engine *>(system); }

int mdrtHipSetPositions(void *system, const double *x) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] {
    e->to_device(e->d_x, x, 3 * e->n * sizeof(double));
    e->forces_valid = false;
  });
}

int mdrtHipSetVelocities(void *system, const double *v) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] {
    e->to_device(e->d_v, v, 3 * e->n * sizeof(double));
  });
}

int mdrtHipGetPositions(void *system, double *x) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] {
    e->to_host(x, e->d_x, 3 * e->n * sizeof(double));
  });
}

int mdrtHipGetVelocities(void *system, double *v) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] {
    e->to_host(v, e->d_v, 3 * e->n * sizeof(double));
  });
}

int mdrtHipComputeForces(void *system, double *forces, double *energies) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->compute(forces, energies); });
}

int mdrtHipComputeForceTerms(void *system, int32_t terms, double *forces, double *energies) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->compute_terms(terms, forces, energies); });
}

int mdrtHipSetPmeTransform(void *system, int32_t mode) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->set_pme_transform(mode); });
}

int mdrtHipSetLaunchShape(void *system, int32_t segment, int32_t bricks, int32_t transform) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->set_launch_shape(segment, bricks, transform); });
}

int mdrtHipGetLaunchShape(void *system, int32_t *shape) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->get_launch_shape(shape); });
}

int mdrtHipAutotune(void *system, int32_t what, int32_t steps, int32_t graphs, double dt,
                    double *ms_per_step) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->autotune(what, steps, graphs, dt, ms_per_step); });
}

int mdrtHipSetLangevin(void *system, double temperature, double friction, uint64_t seed) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->set_langevin(temperature, friction, seed); });
}

int mdrtHipGetBox(void *system, double *box) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] {
    for (int k = 0; k < 3; ++k) box[k] = e->box[k];
  });
}

int mdrtHipScaleBox(void *system, const double *scale) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->scale_box(scale); });
}

int mdrtHipBarostatMove(void *system, double pressure_bar, double temperature, uint64_t seed,
                        int32_t *accepted, double *info) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] {
    const bool keep = e->barostat_move(pressure_bar, temperature, seed, info);
    if (accepted) *accepted = keep;
  });
}

int mdrtHipRunNPT(void *system, int64_t steps, double dt, int32_t graph, int32_t every,
                  double pressure_bar, double temperature, uint64_t seed, double *kinetic) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] {
    e->run_npt(steps, dt, graph, every, pressure_bar, temperature, seed, kinetic);
  });
}

int mdrtHipBarostatStats(void *system, int64_t *counts, double *max_dv) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] {
    if (counts) {
      counts[0] = e->mc_tried;
      counts[1] = e->mc_kept;
    }
    if (max_dv) *max_dv = e->mc_dv;
  });
}

int mdrtHipSetStepCount(void *system, int64_t step) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] {
    if (step < 0) throw mdhip::HipError("the step count must not be negative");
    e->step_count = step;
  });
}

int mdrtHipGetStepCount(void *system, int64_t *step) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { *step = e->step_count; });
}

int mdrtHipSaveState(void *system, double *x, double *v, double *forces, double *list_origin,
                     double *prune_origin, mdrtHipRunState *state) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->save_state(x, v, forces, list_origin, prune_origin, state); });
}

int mdrtHipRestoreState(void *system, const double *x, const double *v,
                        const double *list_origin, const double *prune_origin,
                        const mdrtHipRunState *state) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->restore_state(x, v, list_origin, prune_origin, state); });
}

int mdrtHipRun(void *system, int64_t steps, double dt, int32_t graph, double *kinetic) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->run(steps, dt, graph, kinetic); });
}

int mdrtHipCheckList(void *system, double *forces, double *energies, int64_t *counts) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->check_list(forces, energies, counts); });
}

int mdrtHipListPairs(void *system, double reach, int32_t pruned, int64_t capacity,
                     int32_t *pairs, int64_t *count) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] {
    if (!e->forces_valid || e->list_reach != e->cutoff + e->skin)
      throw mdhip::HipError("no list of a run to read: the positions changed since the last step");
    if (pruned && !(e->prune_skin > 0)) throw mdhip::HipError("the system keeps no pruned list");
    const double far = reach + 3.0e-6 * ((e->box[0] + e->box[1]) + e->box[2]);
    const std::vector<long long> keys =
        pruned ? e->export_pairs(e->d_rows_in, e->d_row_count_in, far, capacity)
               : e->export_pairs(e->d_rows, e->d_row_count, far, capacity);
    const int64_t total = (int64_t)keys.size();
    for (int64_t i = 0; i < std::min(total, capacity); ++i) {
      pairs[2 * i] = (int32_t)(keys[i] >> 32);
      pairs[2 * i + 1] = (int32_t)(keys[i] & 0xffffffff);
    }
    *count = total;
  });
}

int mdrtHipPrepareRun(void *system, int64_t steps, double dt, int32_t graph) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->prepare(steps, dt, graph); });
}

int mdrtHipRunMeasured(void *system, int64_t steps, double dt, int32_t graph,
                       double *pair_ms) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->run(steps, dt, graph, nullptr, pair_ms); });
}

int mdrtHipMinimize(void *system, int64_t steps, double max_step, double *energy) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->minimize(steps, max_step, energy); });
}

int mdrtHipNeighborPairs(void *system, double reach, int32_t excluded, int64_t capacity,
                         int32_t *pairs, int64_t *count) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { *count = e->neighbor_pairs(reach, excluded, capacity, pairs); });
}

int mdrtHipStats(void *system, int64_t *stats) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->stats(stats); });
}

void *mdrtHipFramesCreate(const mdrtHipSystemDesc *desc, int32_t max_frames) {
  FrameBatch *f = new FrameBatch();
  if (guarded([&] { f->create(desc, max_frames); }) != 0) {
    delete f;
    return nullptr;
  }
  return f;
}

void mdrtHipFramesDestroy(void *frames) { delete static_cast<FrameBatch *>(frames); }

int mdrtHipFramesSetParameters(void *frames, const double *charges, const double *sigma,
                               const double *epsilon) {
  FrameBatch *f = static_cast<FrameBatch *>(frames);
  return guarded([&] { f->set_parameters(charges, sigma, epsilon); });
}

int mdrtHipFramesGetParameters(void *frames, double *charges, double *sigma, double *epsilon) {
  FrameBatch *f = static_cast<FrameBatch *>(frames);
  return guarded([&] {
    if (charges) std::copy(f->e.h_q.begin(), f->e.h_q.end(), charges);
    if (sigma) std::copy(f->sig_t.begin(), f->sig_t.end(), sigma);
    if (epsilon) std::copy(f->eps_t.begin(), f->eps_t.end(), epsilon);
  });
}

int mdrtHipFramesEvaluate(void *frames, int32_t count, const double *positions,
                          double *energies, double *forces, double *strain) {
  FrameBatch *f = static_cast<FrameBatch *>(frames);
  return guarded([&] { f->evaluate(count, positions, energies, forces, strain); });
}

int mdrtHipFramesVjp(void *frames, int32_t count, const double *energy_cotangent,
                     const double *position_cotangent, double *d_charges, double *d_sigma,
                     double *d_epsilon) {
  FrameBatch *f = static_cast<FrameBatch *>(frames);
  return guarded([&] {
    f->vjp(count, energy_cotangent, position_cotangent, d_charges, d_sigma, d_epsilon);
  });
}

int mdrtHipFramesVjpPairs(void *frames, int32_t count, const double *energy_cotangent,
                          const double *position_cotangent, int32_t npairs,
                          const int32_t *type_pairs, double *d_sigma, double *d_epsilon) {
  FrameBatch *f = static_cast<FrameBatch *>(frames);
  return guarded([&] {
    f->vjp_pairs(count, energy_cotangent, position_cotangent, npairs, type_pairs, d_sigma,
                 d_epsilon);
  });
}

int mdrtHipFramesSetPmeTransform(void *frames, int32_t mode) {
  FrameBatch *f = static_cast<FrameBatch *>(frames);
  return guarded([&] {
    f->e.set_pme_transform(mode);
    if (f->fm_graph) hipGraphExecDestroy(f->fm_graph);
    f->fm_graph = nullptr;
  });
}

int mdrtHipFramesAutotune(void *frames, int32_t what, int32_t count, const double *positions,
                          const double *reference_forces, int32_t repeats, double *ms_per_call) {
  FrameBatch *f = static_cast<FrameBatch *>(frames);
  return guarded(
      [&] { f->autotune(what, count, positions, reference_forces, repeats, ms_per_call); });
}

int mdrtHipFramesHvp(void *frames, int32_t count, const double *v, double *out) {
  FrameBatch *f = static_cast<FrameBatch *>(frames);
  return guarded([&] { f->hvp(count, v, out); });
}

int mdrtHipFramesForceMatching(void *frames, int32_t count, const double *positions,
                               const double *reference_forces, int32_t graph, double *loss,
                               double *d_charges, double *d_sigma, double *d_epsilon) {
  FrameBatch *f = static_cast<FrameBatch *>(frames);
  return guarded([&] {
    f->force_matching(count, positions, reference_forces, loss, d_charges, d_sigma, d_epsilon,
                      graph);
  });
}

int mdrtHipProfileStep(void *system, int32_t repeats, double *ms) {
  Engine *e = static_cast<Engine *>(system);
  return guarded([&] { e->profile(repeats, ms); });
}

}  // extern "C"

#include "replicas.hip"
#include "noalloc.hip"
#include "trajectory.hip"
