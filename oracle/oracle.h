/* oracle.h — FP64 CPU restatement of the reference's per-step force path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing under mdir_amd/ links, loads or calls
 * this library: it is the checker that tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py compare the HIP path against.
 *
 * Units are the reference's: nm, ps, amu, kJ/mol, e.
 * Every function cites the reference file:line whose semantics it restates.
 */
#ifndef MDIR_ORACLE_H
#define MDIR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Lennard-Jones modifiers (lib/Driver/TermLennardJones.cpp:26-168 and the
   potential switch of lib/Dialect/MD/Transforms/ExpandTruncation.cpp:25-45). */
enum {
  ORC_LJ_NONE = 0,
  ORC_LJ_POTENTIAL_SHIFT = 1,
  ORC_LJ_FORCE_SWITCH = 2,
  ORC_LJ_POTENTIAL_SWITCH = 3
};

/* Energy terms written by orc_nonbonded / orc_total (kJ/mol). */
enum {
  ORC_E_LJ = 0,          /* Lennard-Jones of the non-excluded pairs within rc */
  ORC_E_COUL_DIRECT = 1, /* f qq [erfc(beta r)/r - s] within rc               */
  ORC_E_COUL_EXCL = 2,   /* -f qq erf(beta r)/r over every excluded pair     */
  ORC_E_COUL_RECIP = 3,  /* smooth PME reciprocal sum                        */
  ORC_E_COUL_SELF = 4,   /* -f beta/sqrt(pi) sum q^2  + net-charge background */
  ORC_E_DISP_CORR = 5,   /* dispersion correction                            */
  ORC_E_LJ14 = 6,        /* scaled 1-4 Lennard-Jones                         */
  ORC_E_COUL14 = 7,      /* scaled 1-4 Coulomb                               */
  ORC_E_BOND = 8,        /* (k/2)(r - r0)^2                                  */
  ORC_E_UREY_BRADLEY = 9,/* (k/2)(r13 - r0)^2                                */
  ORC_E_ANGLE = 10,      /* (k/2)(theta - theta0)^2                          */
  ORC_E_DIHEDRAL = 11,   /* k (1 + cos(n phi - phase)), Amber impropers too  */
  ORC_E_IMPROPER = 12,   /* (k/2) wrap(phi - phi0)^2                         */
  ORC_NTERMS = 13
};

/* ---- neighbor matrix: lib/Runtime/Templates/NeighborsMatrix.mlir ---- */
int64_t orc_cell_count(double length, double width);
int64_t orc_cell_range(double length, int64_t count, double reach);
double orc_cell_width(int64_t count, const double box[3], double reach,
                      double least);
int64_t orc_neighbors_matrix(int64_t n, const double *x, const double box[3],
                             double reach, double cell_width, int32_t *counts,
                             int32_t *index, int64_t row_width);
void orc_exclusion_filter(int64_t n, int32_t *counts, int32_t *index,
                          int64_t row_width, const int32_t *excl_start,
                          const int32_t *excl_atoms);

/* ---- nonbonded pairs (TermNonbonded.cpp, TermLennardJones.cpp) ---- */
/* sigma/epsilon are ntypes x ntypes tables.  excl_* is a symmetric CSR
   list of excluded partners.  Forces are ADDED to F (n x 3).  energies has
   ORC_NTERMS entries and is ADDED to.  mixed != 0 evaluates the pair
   arithmetic in f32 (the reference's mixed mode) for timing only.
   virial (9, row-major, may be NULL) is ADDED to: W_ab = sum r_a F_b. */
int orc_nonbonded(int64_t n, const double *x, const double *q,
                  const int32_t *type, int32_t ntypes, const double *sigma,
                  const double *epsilon, const double box[3], double cutoff,
                  double switch_from, int32_t lj_modifier,
                  int32_t coulomb_shift, double beta, double coulomb,
                  const int32_t *excl_start, const int32_t *excl_atoms,
                  int32_t mixed, double *F, double *energies, double *virial);

/* Excluded-pair Ewald correction over a list of pairs (i,j), no cutoff
   (TermNonbonded.cpp:1615-1671). */
void orc_excluded(int64_t npairs, const int32_t *pairs, const double *x,
                  const double *q, const double box[3], double beta,
                  double coulomb, double *F, double *energy);

/* Scaled 1-4 pairs (TermPairs14.cpp:11-17, :97-99): LJ with the pair's
   sigma and eps*scale_lj under the same modifier, plus f qq scale_c / r. */
void orc_pairs14(int64_t npairs, const int32_t *pairs, const double *sigma,
                 const double *eps_scaled, const double *qq_scaled,
                 const double *x, const double box[3], double cutoff,
                 double switch_from, int32_t lj_modifier, double *F,
                 double *e_lj, double *e_coul);

/* Ewald constants (TermNonbonded.cpp:573-636, :651; 05-electrostatics.md:39-40). */
double orc_ewald_beta(double cutoff, double tolerance);
int64_t orc_smooth_size(int64_t least);
double orc_self_energy(int64_t n, const double *q, double beta, double coulomb,
                       const double box[3]);
double orc_dispersion_energy(int64_t n, const int32_t *type, int32_t ntypes,
                             const double *sigma, const double *epsilon,
                             int64_t nexcl_pairs, const int32_t *excl_pairs,
                             const double box[3], double cutoff);

/* ---- PME: PMECode.h, Influence.cpp, HostPME*.cpp ---- */
void orc_bspline(double w, int order, double *M, double *dM);
void orc_spline_moduli(int64_t count, int order, double *moduli);
void orc_pme_spread(int64_t n, const double *x, const double *q,
                    const double box[3], const int64_t K[3], int order,
                    double *grid);
void orc_pme_influence(const double box[3], const int64_t K[3], int order,
                       double beta, double coulomb, double *bc);
void orc_pme_gather(int64_t n, const double *x, const double *q,
                    const double box[3], const int64_t K[3], int order,
                    const double *phi, double *F);
/* Full chain: spread, R2C, energy/virial, convolution, C2R, gather.
   F is ADDED to.  virial[6] = xx yy zz xy xz yz (may be NULL). */
double orc_pme(int64_t n, const double *x, const double *q,
               const double box[3], const int64_t K[3], int order,
               double beta, double coulomb, double *F, double *virial);

/* ---- FFT: runtime/mdrt_fft.c conventions ---- */
/* In-place complex line transform, sign -1 forward, +1 backward,
   unnormalised.  data = n interleaved (re, im). */
void orc_fft_line(int64_t n, double *data, int sign);
void orc_dft_line(int64_t n, const double *in, double *out, int sign);
/* real k1 x k2 x k3 -> complex k1 x k2 x (k3/2+1) interleaved, and back
   (the backward overwrites nothing it reads: complex is copied). */
void orc_fft3d_r2c(const int64_t K[3], const double *real, double *cplx);
void orc_fft3d_c2r(const int64_t K[3], const double *cplx, double *real);

/* Ewald reciprocal sum over wave vectors until the Gaussian is 1e-18
   (method of test/Driver/Inputs/ewald_reference.py:99-117).  F may be NULL. */
double orc_ewald_reciprocal(int64_t n, const double *x, const double *q,
                            const double box[3], double beta, double coulomb,
                            double *F);

/* ---- constraints and integration (ConstraintIR.cpp, DynamicsIR.cpp) ---- */
/* SETTLE of nw waters whose atoms are (o, o+1, o+2) for o in first[]. */
void orc_settle_positions(int64_t nw, const int32_t *first, double m_o,
                          double m_h, double d_oh, double d_hh,
                          const double *x_old, double *x_new);
void orc_settle_velocities(int64_t nw, const int32_t *first, double m_o,
                           double m_h, const double *x, double *v);
/* Velocity-Verlet NVE of the full nonbonded + PME system with SETTLE
   waters.  Runs `steps` steps; energies per step of (potential, kinetic)
   written to trace (2*steps, may be NULL). */
typedef struct {
  int64_t n;
  const double *q;
  const int32_t *type;
  int32_t ntypes;
  const double *sigma, *epsilon;
  const double *mass;
  double box[3];
  double cutoff, switch_from;
  int32_t lj_modifier, coulomb_shift, dispersion;
  double beta, coulomb;
  const int32_t *excl_start, *excl_atoms;
  int64_t nexcl_pairs;
  const int32_t *excl_pairs;
  int64_t K[3];
  int32_t order;
  int64_t nwaters;
  const int32_t *water_first;
  double m_o, m_h, d_oh, d_hh;
  /* 1-4 pairs (orc_pairs14) */
  int64_t npairs14;
  const int32_t *pairs14;
  const double *p14_sigma, *p14_eps, *p14_qq;
  /* bonded tuples (orc_bonded): 4 atoms, a kind, 4 parameters each */
  int64_t ntuples;
  const int32_t *tuple_atoms, *tuple_kind;
  const double *tuple_param;
  /* constraint clusters (orc_shake_positions): a centre atom and the atoms
     bonded to it at fixed lengths, atoms first .. first + size - 1 */
  int64_t nclusters;
  const int32_t *cluster_first, *cluster_size, *cluster_center;
  const double *cluster_length; /* per atom: its length to the centre */
} orc_system;

/* Bonded tuples (lib/Driver/TermBond.cpp, TermUreyBradley.cpp,
   TermAngle.cpp, TermDihedral.cpp, TermImproper.cpp; the coordinates and
   their gradients of lib/Dialect/MD/MDCoordinates.cpp:279-400).  kind: 0
   bond (i, j; k, r0), 1 Urey-Bradley (i, k; k, r0), 2 angle (i, j, k apex
   j; k, theta0), 3 dihedral (i, j, k, l; k, n, phase), 4 harmonic improper
   (i, j, k, l; k, phi0).  Energies are ADDED to e[ORC_E_BOND ..], forces to
   F. */
void orc_bonded(int64_t ntuples, const int32_t *atoms, const int32_t *kind,
                const double *param, const double *x, const double box[3], double *F,
                double *energies);
/* M-SHAKE of the constraint clusters (ConstraintIR.cpp:595-747): six
   Newton sweeps of the linearized bond equations, the members moving along
   the old bonds; then the velocity constraint, one linear solve. */
void orc_shake_positions(const orc_system *s, const double *x_old, double *x_new);
void orc_rattle_velocities(const orc_system *s, const double *x, double *v);

int orc_total(const orc_system *s, const double *x, double *F,
              double *energies, int32_t mixed);
/* The CPU baseline: velocity Verlet with SETTLE, the neighbor matrix kept
   until an atom moved half the buffer, f32 pair rows, fixed-point PME. */
typedef struct orc_md orc_md;
orc_md *orc_md_create(const orc_system *s, double skin);
void orc_md_destroy(orc_md *c);
int orc_md_run(orc_md *c, double *x, double *v, double dt, int64_t steps);
int64_t orc_md_builds(const orc_md *c);
void orc_set_threads(int threads);
/* Langevin dynamics (g-BAOAB) and its Philox 4x32-10 noise. */
void orc_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]);
void orc_langevin_normals(uint64_t seed, int64_t step, uint32_t id, double r[3]);
int orc_langevin(const orc_system *s, double *x, double *v, double dt, int64_t steps,
                 int64_t first_step, uint64_t seed, double gamma, double kT,
                 double *trace, int32_t mixed);
/* mixed != 0 evaluates the pair arithmetic in f32 (timing only). */
int orc_nve(const orc_system *s, double *x, double *v, double dt,
            int64_t steps, double *trace, int32_t mixed);
int orc_threads(void);

/* ---- initial velocities (System.cpp:1298-1350; InitialVelocityDraws.h) ---- */
uint64_t orc_splitmix64(uint64_t *state);
uint64_t orc_xoshiro256ss(uint64_t state[4]);
void orc_initial_normals(uint64_t seed, int64_t n, double *out);
void orc_initial_velocities(uint64_t seed, int64_t n, const double *mass,
                            double temperature, const double *x,
                            int64_t nwaters, const int32_t *water_first,
                            int64_t dof, double *v);

#ifdef __cplusplus
}
#endif
#endif
