/* mdrt_hip.h — C ABI of libmdrt_hip.so, the MI355X (gfx950) runtime and
 * per-step force path.
 *
 * Groups of entry points:
 *
 *  1. The runtime ABI of the reference's runtime/mdrt_cuda.c, same names and
 *     signatures, on HIP (modules are gfx950 code objects, streams are HIP
 *     streams, the FFT entry points run this library's own kernels).
 *
 *  2. The force path as explicit entry points (the reference emits these
 *     kernels from lib/Conversion/MDExecToGPU; a compiled program cannot be
 *     produced without its MLIR toolchain, so the path is exposed here with
 *     plain pointers and sizes): a system handle, forces and energies,
 *     velocity-Verlet steps with SETTLE waters, and the neighbor pair set.
 *
 *  3. The frame evaluator: energies, forces and reverse products (charges,
 *     per-type sigma and epsilon; position Hessian products) of many frames
 *     of one system per launch, and force matching in one call.
 *
 *  4. Replica sets: many systems of uneven size stepped together.
 *
 *  5. For callers that cannot allocate device memory.
 *
 *  6. Outputs: trajectories in DCD, written from staged copies on a thread
 *     of their own (the reference's driver, lib/Driver/Trajectory.cpp and
 *     lib/Driver/Output.cpp, over group 1's staging entry points).
 *
 * Units are the reference's: nm, ps, amu, kJ/mol, e.  All pointers in
 * group 2 are host pointers unless the name says device.
 */
#ifndef MDRT_HIP_H
#define MDRT_HIP_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------
 * 1. Runtime ABI (runtime/mdrt_cuda.c)
 * ------------------------------------------------------------------- */

typedef void *mgpuModule;   /* hipModule_t   */
typedef void *mgpuFunction; /* hipFunction_t */
typedef void *mgpuStream;   /* hipStream_t   */
typedef void *mgpuEvent;    /* hipEvent_t    */

/* mdrt_cuda.c:420 */
void mgpuSetDefaultDevice(int32_t device);
/* mdrt_cuda.c:476 — data is a gfx950 code object (ELF) */
mgpuModule mgpuModuleLoad(void *data, size_t size);
/* mdrt_cuda.c:502 — no JIT on this path: a code object is loaded as is */
mgpuModule mgpuModuleLoadJIT(void *data, int optLevel, size_t size);
/* mdrt_cuda.c:526 */
void mgpuModuleUnload(mgpuModule module);
/* mdrt_cuda.c:553 */
mgpuFunction mgpuModuleGetFunction(mgpuModule module, const char *name);
/* mdrt_cuda.c:669 */
void mgpuLaunchKernel(mgpuFunction function, intptr_t gridX, intptr_t gridY,
                      intptr_t gridZ, intptr_t blockX, intptr_t blockY,
                      intptr_t blockZ, int32_t smem, mgpuStream stream,
                      void **params, void **extra, size_t paramsCount);
/* mdrt_cuda.c:679 — CDNA has no thread-block clusters: cluster extents
   other than 1 are refused. */
void mgpuLaunchKernelCooperative(mgpuFunction function, intptr_t gridX,
                                 intptr_t gridY, intptr_t gridZ,
                                 intptr_t clusterX, intptr_t clusterY,
                                 intptr_t clusterZ, intptr_t blockX,
                                 intptr_t blockY, intptr_t blockZ,
                                 int32_t smem, mgpuStream stream,
                                 void **params, void **extra);
/* mdrt_cuda.c:401 */
intptr_t mdrtCooperativeGridSize(intptr_t requested);
/* mdrt_cuda.c:695, :775, :1747, :1757 */
mgpuStream mgpuStreamCreate(void);
void mgpuStreamDestroy(mgpuStream stream);
void mgpuStreamSynchronize(mgpuStream stream);
void mgpuStreamWaitEvent(mgpuStream stream, mgpuEvent event);
/* mdrt_cuda.c:1762-1782 */
mgpuEvent mgpuEventCreate(void);
void mgpuEventDestroy(mgpuEvent event);
void mgpuEventSynchronize(mgpuEvent event);
void mgpuEventRecord(mgpuEvent event, mgpuStream stream);
/* mdrt_cuda.c:2147, :2238 */
void *mgpuMemAlloc(uint64_t size, mgpuStream stream, bool isHostShared);
void mgpuMemFree(void *pointer, mgpuStream stream);
/* mdrt_cuda.c:2431 — finished on return when either end is host memory */
void mgpuMemcpy(void *destination, void *source, size_t size,
                mgpuStream stream);
/* mdrt_cuda.c:2471, :2490 — a copy on the current stream, finished on
   return without waiting for work on the side streams; and the same wait
   without a copy */
void mdrtCudaMemcpyScoped(uintptr_t destination, uintptr_t source,
                          uint64_t size);
void mdrtCudaWaitScoped(void);
/* mdrt_cuda.c:2494, :2510 */
void mgpuMemset32(void *destination, unsigned int value, size_t count,
                  mgpuStream stream);
void mgpuMemset16(void *destination, unsigned short value, size_t count,
                  mgpuStream stream);
/* mdrt_cuda.c:729-742 — side streams joined by events */
void mdrtSideBegin(void);
void mdrtSideEnd(void);
void mdrtSide2Begin(void);
void mdrtSide2End(void);
void mdrtSideJoin(void);
/* mdrt_cuda.c:762, :770 */
void *mdrtUseStream(void *stream);
void mdrtRestoreStream(void *previous);

/* The descriptor of a memref of one dimension (mdrt_cuda.c:1530-1536). */
typedef struct {
  void *allocated;
  void *aligned;
  int64_t offset;
  int64_t size;
  int64_t stride;
} mdrtBuffer1;

/* mdrt_cuda.c:1615, :1642, :1691, :1710 — flags in mapped host memory */
void _mlir_ciface_mdrtFlagMemory(mdrtBuffer1 *result, int64_t place);
int64_t mdrtFlagStart(void *device);
int32_t mdrtFlagFinish(int64_t slot);
int32_t mdrtFlagTake(void *device);
/* mdrt_cuda.c:1548 */
void _mlir_ciface_mdrtHostResults(mdrtBuffer1 *result, int64_t place,
                                  int64_t count);
/* mdrt_cuda.c:154 */
void mdrtCountHostRead(size_t bytes, int waits, int flags);
/* mdrt_cuda.c:2566 */
void mdrtCudaSetStopHandler(void (*handler)(const char *));
/* mdrt_cuda.c:3817 — called before the runtime aborts on a failed device
   call or a misuse of the ABI */
void mdrtSetFailureHook(void (*hook)(void));
/* mdrt_cuda.c:2824, :2836 */
uint64_t mdrtDeviceTrim(void);
void mdrtDeviceMemoryCounts(int64_t counts[8]);

/* mdrt_cuda.c:3530-3576 — kept memory: device memory outside the pool that
   an embedding program holds across runs, and copies to and from it on the
   current stream (CopyWithin does not wait; CopyToHost returns when the
   bytes are on the host). */
void *mdrtDeviceAllocateKept(uint64_t size);
void mdrtDeviceFreeKept(void *pointer);
void mdrtDeviceCopyWithin(void *destination, const void *source, uint64_t size);
void mdrtDeviceCopyToHost(void *destination, const void *source, uint64_t size);

/* mdrt_cuda.c:3611-3845 — staged copies of a frame while the steps go on.
   A slot holds up to 4 parts.  Copy: a device-to-device copy of `size`
   bytes at `source` into the slot's device buffer of that part, on the
   current stream; returns the pinned host memory the bytes will be in.
   End: a stream of copies of the runtime's own waits for the current
   stream, copies the parts copied since the last end to the host, and then
   has the device store the end's serial number into the slot's mapped
   word.  Wait (any thread): 0 once the copies of the last end are on the
   host, else the HIP error of the wait, with the call that returned it in
   *call, which the thread of the steps hands to Fail (it reports the error
   as any failed device call is reported, and aborts).  The environment
   variable MDRT_STAGING_MARK_DELAY=<ms> holds the stored word back by that
   long (a host function on the stream of copies), for tests of a word that
   lands after its wait has returned. */
void *mdrtStagingCreate(void);
void *mdrtStagingCopy(void *slot, int32_t part, const void *source, uint64_t size);
void mdrtStagingEnd(void *slot);
int32_t mdrtStagingWait(void *slot, const char **call);
void mdrtStagingFail(int32_t result, const char *call);
void mdrtStagingDestroy(void *slot);

/* mdrt_cuda.c:2901-2991 — the 3D transforms of the PME grid, with this
   library's FFT kernels: real k1 x k2 x k3 <-> half spectrum
   k1 x k2 x (k3/2+1) interleaved complex, forward exp(-2 pi i),
   backward exp(+2 pi i), unnormalised.  Buffers are device memory. */
void _mlir_ciface_mdrtCudaFFTForward3D(mdrtBuffer1 *real, mdrtBuffer1 *complex,
                                       int64_t k1, int64_t k2, int64_t k3);
void _mlir_ciface_mdrtCudaFFTBackward3D(mdrtBuffer1 *complex,
                                        mdrtBuffer1 *real, int64_t k1,
                                        int64_t k2, int64_t k3);
void _mlir_ciface_mdrtCudaFFTForward3DF32(mdrtBuffer1 *real,
                                          mdrtBuffer1 *complex, int64_t k1,
                                          int64_t k2, int64_t k3);
void _mlir_ciface_mdrtCudaFFTBackward3DF32(mdrtBuffer1 *complex,
                                           mdrtBuffer1 *real, int64_t k1,
                                           int64_t k2, int64_t k3);
/* mdrt_cuda.c:2986 — backward with the separable Gaussian influence applied
   as the transform loads the spectrum (pme_callback.cu:5-21): tables holds
   rows t_1 | t_2 | t_3 (t_1 includes f/(pi V)) then m_1 | m_2 | m_3, each
   row max(k1,k2,k3) long. */
void _mlir_ciface_mdrtCudaFFTGaussianBackwardF32(mdrtBuffer1 *complex,
                                                 mdrtBuffer1 *real,
                                                 mdrtBuffer1 *tables,
                                                 int64_t k1, int64_t k2,
                                                 int64_t k3);
/* mdrt_cuda.c:2765, :2777 */
int mdrtRetainFFTPlan(int64_t k1, int64_t k2, int64_t k3, int forward,
                      int single, const float *tables);
void mdrtReleaseFFTPlan(int handle);

/* ---------------------------------------------------------------------
 * 2. The force path
 * ------------------------------------------------------------------- */

enum {
  MDRT_LJ_NONE = 0,            /* TermLennardJones.cpp:66-81            */
  MDRT_LJ_POTENTIAL_SHIFT = 1, /* TermLennardJones.cpp:108-111          */
  MDRT_LJ_FORCE_SWITCH = 2,    /* TermLennardJones.cpp:112-135          */
  MDRT_LJ_POTENTIAL_SWITCH = 3 /* ExpandTruncation.cpp:19-45            */
};

/* Energy terms, kJ/mol. */
enum {
  MDRT_E_LJ = 0,
  MDRT_E_COUL_DIRECT = 1,
  MDRT_E_COUL_EXCL = 2,
  MDRT_E_COUL_RECIP = 3,
  MDRT_E_COUL_SELF = 4,
  MDRT_E_DISP_CORR = 5,
  MDRT_E_LJ14 = 6,
  MDRT_E_COUL14 = 7,
  MDRT_E_KINETIC = 8,
  MDRT_E_BONDS = 9,           /* TermBond.cpp          (k/2)(r - r0)^2          */
  MDRT_E_UREY_BRADLEY = 10,   /* TermUreyBradley.cpp   (k/2)(r13 - r0)^2        */
  MDRT_E_ANGLES = 11,         /* TermAngle.cpp         (k/2)(theta - theta0)^2  */
  MDRT_E_DIHEDRALS = 12,      /* TermDihedral.cpp      k (1 + cos(n phi - d))   */
  MDRT_E_IMPROPERS = 13,      /* TermImproper.cpp      (k/2) wrap(phi - phi0)^2 */
  MDRT_NUM_ENERGIES = 14
};

/* Kinds of bonded tuples (the atoms they take; their four parameters). */
enum {
  MDRT_TUPLE_BOND = 0,          /* i, j;       k, r0          */
  MDRT_TUPLE_UREY_BRADLEY = 1,  /* i, k;       k, r0          */
  MDRT_TUPLE_ANGLE = 2,         /* i, j, k (apex j); k, theta0 */
  MDRT_TUPLE_DIHEDRAL = 3,      /* i, j, k, l; k, n, phase    */
  MDRT_TUPLE_IMPROPER = 4       /* i, j, k, l; k, phi0        */
};

/* A system: the topology and settings of lib/Driver (Topology, Control).
   Arrays are copied at creation. */
typedef struct {
  int64_t n;                  /* particles                              */
  const double *positions;    /* n x 3                                  */
  const double *velocities;   /* n x 3 or NULL (zero)                   */
  const double *masses;       /* n                                      */
  const double *charges;      /* n                                      */
  const int32_t *types;       /* n, in [0, ntypes)                      */
  int32_t ntypes;             /* <= 32                                  */
  const double *sigma;        /* ntypes x ntypes                        */
  const double *epsilon;      /* ntypes x ntypes                        */
  double box[3];              /* orthorhombic, each edge > 2 (rc+skin)  */
  double cutoff;              /* rc                                     */
  double switch_from;         /* rs for the switches                    */
  int32_t lj_modifier;        /* MDRT_LJ_*                              */
  int32_t coulomb_shift;      /* POTENTIAL_SHIFT of the direct Coulomb  */
  int32_t dispersion;         /* dispersion correction on/off           */
  double beta;                /* Ewald beta (nm^-1); 0: no Coulomb      */
  int32_t pme_grid[3];        /* K1 K2 K3 (factors 2,3,5,7)             */
  int32_t pme_order;          /* 4, 6 or 8                              */
  int64_t nexclusions;
  const int32_t *exclusions;  /* nexclusions x 2                        */
  int64_t npairs14;
  const int32_t *pairs14;     /* npairs14 x 2                           */
  const double *pairs14_sigma;
  const double *pairs14_eps;  /* eps * s_LJ                             */
  const double *pairs14_qq;   /* f q_i q_j s_C                          */
  int64_t nwaters;
  const int32_t *water_first; /* first atom (oxygen) of each water      */
  double d_oh, d_hh;          /* SETTLE geometry                        */
  double skin;                /* list buffer beyond rc (nm)             */
  double prune_skin;          /* buffer of the pruned list the pair loop
                                 runs over (nm, < skin); 0: no pruning   */
  /* bonded tuples (lib/Driver/TermBond.cpp, TermAngle.cpp, ...): atoms 4
     each (unused 0), a kind MDRT_TUPLE_*, parameters 4 each */
  int64_t ntuples;
  const int32_t *tuple_atoms;
  const int32_t *tuple_kind;
  const double *tuple_param;
  /* bond constraints in clusters (ConstraintIR.cpp:595-747, M-SHAKE): the
     atoms first .. first + size - 1 (size <= 5), the centre at
     first + centre, the others bonded to it at cluster_length[atom] */
  int64_t nclusters;
  const int32_t *cluster_first;
  const int32_t *cluster_size;
  const int32_t *cluster_center;
  const double *cluster_length; /* n (0 for atoms in no cluster)         */
} mdrtHipSystemDesc;

/* Creates a system on the current device; NULL on error (see
   mdrtHipLastError). */
void *mdrtHipSystemCreate(const mdrtHipSystemDesc *desc);
void mdrtHipSystemDestroy(void *system);
const char *mdrtHipLastError(void);

int mdrtHipSetPositions(void *system, const double *x);
int mdrtHipSetVelocities(void *system, const double *v);
int mdrtHipGetPositions(void *system, double *x);
int mdrtHipGetVelocities(void *system, double *v);

/* Forces (n x 3, may be NULL) and energies (MDRT_NUM_ENERGIES, may be NULL)
   at the current positions; builds the list. */
int mdrtHipComputeForces(void *system, double *forces, double *energies);

/* Force terms of mdrtHipComputeForceTerms. */
enum {
  MDRT_TERM_PAIRS = 1,    /* nonbonded pairs within rc: LJ + direct Coulomb */
  MDRT_TERM_SPECIAL = 2,  /* excluded-pair Ewald correction and 1-4 pairs */
  MDRT_TERM_RECIP = 4,    /* PME reciprocal sum */
  MDRT_TERM_BONDED = 8    /* bonded tuples: bonds, angles, dihedrals, ... */
};

/* As mdrtHipComputeForces, for the terms of the mask `terms` only (the
   reference states its precision per force term, docs/paper/07-precision.md;
   the per-term split of TermNonbonded.cpp's nonbonded, excluded and
   reciprocal terms); the energies of the other terms are 0. */
int mdrtHipComputeForceTerms(void *system, int32_t terms, double *forces, double *energies);

/* The y and x transforms of the PME planes: 0 = the LDS Stockham FFT (the
   default), 1 = DFT matrix products on the matrix cores in split BF16x3
   (f32 operands as three bf16 pieces, six MFMA products per real product,
   f32 accumulation): whole planes in LDS for plane edges of 32 or 64, and
   for any other x and y edges up to 224 tiles of 32 lines padded to a
   multiple of 32 (the spread and gather then write and read the half
   spectrum as [x][y][kz]); wider grids are refused, and frame batches take
   edges of 32 or 64 only.  Replaces the reference's cuFFT plan choice
   (mdrt_cuda.c:2646-2991) and axis transform (AxisFFTEmitter.cpp:305) with
   this library's own. */
int mdrtHipSetPmeTransform(void *system, int32_t mode);

/* One launch shape, set directly (as the trial below would keep it):
   pair-loop segments of 16, 24 or 48 row entries, PME bricks of 8, 4, 2 or
   1 complex z lines (those that fit LDS at this grid), plane transform 0 or
   1 (as mdrtHipSetPmeTransform). The list is rebuilt at the next
   evaluation. The counterpart of a launch shape chosen in the reference's
   TOWERS set-up (tools/mdir/TowersShape.h). */
int mdrtHipSetLaunchShape(void *system, int32_t segment, int32_t bricks, int32_t transform);

/* The current shape: shape[3] = {segment entries, brick z lines (0 without
   PME), transform}. */
int mdrtHipGetLaunchShape(void *system, int32_t *shape);

/* Launch-shape trial: times the launch shapes of this system's kernels
   and keeps the fastest, replacing the reference's TOWERS launch-shape
   trial and its step-graph trial (tools/mdir/TowersTrial.h:13-33,
   Run.cpp:913-926). Bit 0 of `what`: PME bricks of 8, 4, 2 or 1 complex z
   lines (those that fit LDS). Bit 1: the plane transform (0 or 1 of
   mdrtHipSetPmeTransform, where the grid allows 1). Bit 2: pair-loop
   segments of 16, 24 or 48 row entries per wave. Each candidate runs one
   untimed captured graph of `steps` steps of dt (ps), then `graphs` timed
   ones, from the same positions and velocities, which are restored at the
   end; two passes over the candidates, the faster time of each counting,
   and the current shape stays unless another is 2 % faster.
   ms_per_step[24]: [8 * s + 4 * transform + k] for segments of 16, 24, 48
   (s = 0, 1, 2) and bricks of TL = 8 >> k, NaN for shapes not run. The
   shapes give the same forces up to the f32 rounding of the gather's
   per-brick partial sums (up to about 1e-7 relative). */
int mdrtHipAutotune(void *system, int32_t what, int32_t steps, int32_t graphs, double dt,
                    double *ms_per_step);

/* Langevin dynamics for the steps of mdrtHipRun: friction gamma (1/ps) > 0
   makes each step the g-BAOAB step (kick, half drift, O, half drift, kick;
   SETTLE and the velocity constraint after each part) that the reference's
   Langevin integrator with constraints takes (lib/Driver/DynamicsIR.cpp,
   geodesic branch; docs/dialects/dyn.md:441-460 for the O step):
   v <- exp(-gamma dt) v + sqrt((1 - exp(-2 gamma dt)) kB T / m) R, R three
   normals per atom from one Philox 4x32-10 block of (seed, step number,
   atom index, stream 2), as runtime/mdrt.c:289-312 and
   lib/Driver/NoiseIR.cpp draw them. friction 0 restores velocity Verlet. */
int mdrtHipSetLangevin(void *system, double temperature, double friction, uint64_t seed);

/* The box edges (nm). */
int mdrtHipGetBox(void *system, double *box /* 3 */);

/* Scales every unit's centre of mass (a SETTLE water, or a free atom) by
   scale[a] along axis a and the box with it; units keep their shape. */
int mdrtHipScaleBox(void *system, const double *scale /* 3 */);

/* One isotropic Monte Carlo volume move at pressure_bar and temperature (K)
   (the NPT coupling of the reference's Amber comparison runs,
   scripts/benchmarks/amber/README.md:56, with the reference's pressure
   units, bar): a volume change uniform in [-dv, dv] (dv adapted towards
   25-75 % acceptance), units' centres scaled, the energy evaluated at the
   new volume, and the move kept with probability
   min(1, exp(-W / kT)), W = U1 - U0 + P dV - N_units kT ln(V1 / V0).
   Random numbers: Philox 4x32-10 of (seed, step number, 0, stream 1).
   *accepted (may be NULL) 1 if kept; info[4] (may be NULL): U0, U1, dV
   (nm^3), W (kJ/mol). */
int mdrtHipBarostatMove(void *system, double pressure_bar, double temperature, uint64_t seed,
                        int32_t *accepted, double *info);

/* `steps` steps of mdrtHipRun (Langevin when set) with a volume move after
   every `every` steps. */
int mdrtHipRunNPT(void *system, int64_t steps, double dt, int32_t graph, int32_t every,
                  double pressure_bar, double temperature, uint64_t seed, double *kinetic);

/* Volume moves tried and kept (counts[2]) and the current largest volume
   change (nm^3). */
int mdrtHipBarostatStats(void *system, int64_t *counts, double *max_dv);

/* The step number of the next step (the Langevin noise's step counter):
   0 at creation, advanced by every step mdrtHipRun takes (a restart sets it
   to the checkpoint's count). */
int mdrtHipSetStepCount(void *system, int64_t step);
int mdrtHipGetStepCount(void *system, int64_t *step);

/* The state of a run beyond its positions and velocities, as a checkpoint
   keeps it (the reference's include/mdir/Driver/Checkpoint.h:36-136 holds
   the same quantities in H5MD, lib/Driver/Checkpoint.cpp). */
typedef struct {
  int64_t step_count;          /* the step number of the next step        */
  double box[3];               /* nm                                      */
  double barostat_dv;          /* the volume moves' largest change, nm^3  */
  int64_t barostat_counts[4];  /* moves tried, kept; in the current
                                  window of 10: tried, kept              */
  int32_t list_kept;           /* 1: the origins are those of the run's
                                  list; 0: the next run builds one at x  */
  int32_t reserved;
} mdrtHipRunState;

/* The run's state: positions, velocities, the forces the next step begins
   with (f64, atom order; NULL to skip), the positions the current list was
   built at and those it was last pruned at (n x 3 each, or NULL), and
   *state.  Restore puts such a state back: it builds the list where the
   saved run built it (and prunes where it pruned), then evaluates the
   forces at `x` through that list, so the steps that follow are those the
   saved run takes, bit for bit. */
int mdrtHipSaveState(void *system, double *x, double *v, double *forces, double *list_origin,
                     double *prune_origin, mdrtHipRunState *state);
int mdrtHipRestoreState(void *system, const double *x, const double *v,
                        const double *list_origin, const double *prune_origin,
                        const mdrtHipRunState *state);

/* `steps` velocity-Verlet steps of dt (ps) with SETTLE waters.  graph != 0
   replays the step sequence as captured HIP graphs of `graph` steps (and
   one of the remainder steps mod graph).
   kinetic (may be NULL) receives the kinetic energy after the last step. */
int mdrtHipRun(void *system, int64_t steps, double dt, int32_t graph,
               double *kinetic);

/* The check of the list between rebuilds, the counterpart of the
   reference's MDIR_CHECK_PRUNING (runtime/mdrt.c:48-79, mdrtCheckPruned):
   after mdrtHipRun, the forces (n x 3, may be NULL) and energies (may be
   NULL) the next step would compute at the current positions, through the
   list the steps hold (no forced rebuild: the build runs only if the last
   drift asked for one, as in a step).  counts[3] (may be NULL): the pairs
   within rc (+ 3e-6 (Lx+Ly+Lz), the pair-set margin) found in the list and
   in the rows the pair loop runs over (the pruned rows with a dual list;
   the two are equal unless the pruned list misses a pair), and the list
   builds the check itself did.  Fails when the positions changed since
   the last step. */
int mdrtHipCheckList(void *system, double *forces, double *energies, int64_t *counts);

/* The pairs (i < j, excluded pairs dropped) of the list the steps hold
   (pruned != 0: the pruned rows the pair loop runs over) within
   reach + 3e-6 (Lx+Ly+Lz) at the current positions, by the f32 test of the
   reference's matrix (NeighborsMatrix.mlir:217-232), without a rebuild:
   the pair set behind mdrtHipCheckList's counts.  Writes at most
   `capacity` pairs; *count receives the number found.  Fails when the
   positions changed since the last step. */
int mdrtHipListPairs(void *system, double reach, int32_t pruned, int64_t capacity,
                     int32_t *pairs, int64_t *count);

/* Captures, without running them, the step graphs a later
   mdrtHipRun(system, steps, dt, graph) replays: the graph of `graph` steps
   and the graph of the remainder steps mod graph (each run replays only
   captured graphs when graph != 0; this keeps the capture out of a timed
   run).  No counterpart in the reference, whose step graph is captured by
   mdrtStepCapture (runtime/mdrt_step_graph.cu:158-464) inside the run. */
int mdrtHipPrepareRun(void *system, int64_t steps, double dt, int32_t graph);

/* mdrtHipRun without the kinetic energy, timing the pair loop live: device
   clock (wall_clock64) readings by one-wave kernels just before and after
   the pair kernel of every step, and once more right after, on its stream,
   inside the captured graphs too; *pair_ms receives the mean of
   (after - before) - (again - after) (ms).  No counterpart in the
   reference: it is the measurement hook of the benchmark. */
int mdrtHipRunMeasured(void *system, int64_t steps, double dt, int32_t graph,
                       double *pair_ms);

/* Steepest descent of `steps` accepted-or-rejected moves, rigid waters. */
int mdrtHipMinimize(void *system, int64_t steps, double max_step,
                    double *energy);

/* The neighbor pair set of the reference's matrix build
   (NeighborsMatrix.mlir:193-460) at `reach`: unordered pairs (i < j) whose
   f32 minimum-image distance is within reach + 3e-6 (Lx+Ly+Lz), read from
   this library's cluster list.  With `excluded` != 0 the excluded pairs are
   dropped as the exclusion filter of MDExecKernels.cpp:1748-1823 drops them.
   Writes at most `capacity` pairs; *count receives the number found
   (duplicates removed). */
int mdrtHipNeighborPairs(void *system, double reach, int32_t excluded,
                         int64_t capacity, int32_t *pairs, int64_t *count);

/* List statistics: places, clusters, cluster-pair tiles of the list,
   largest group row (entries), row capacity, columns, column capacity,
   builds done since creation, tiles of the rows the pair loop runs over
   (the pruned list when pruning), prunes done since creation. */
int mdrtHipStats(void *system, int64_t *stats /* 10 */);

/* Device-event timing of each kernel of a step on the stream it runs on,
   ms per launch averaged over `repeats` steps (the steps advance the
   state): integrate_first, nb_bin, nb_sort, nb_search, nb_pairs (with the
   excluded and 1-4 pairs), pme_bin (atoms listed per brick),
   pme_spread_rows (spread + forward z transform), pme_yx (y and x transforms
   with the influence product), pme_gather_rows (backward z transform +
   gather), integrate_second. */
int mdrtHipProfileStep(void *system, int32_t repeats, double *ms /* 10 */);

/* ---------------------------------------------------------------------
 * 3. Frame evaluator: energies, forces and reverse products of many frames
 *    of one system (shared topology and box) in one launch per kernel.
 *
 *    Replaces the device frame batch of the reference's compiled entry
 *    (lib/Compiler/Simulation.cpp:3189 evaluateFrames; per-frame body
 *    lib/Driver/FrameBatchEvaluationIR.cpp:20-60) behind the Python
 *    FrameEvaluator (python/mdir/_frames.py:458, evaluate :715,
 *    FrameEnergies.vjp :327) and the torch operators mdir::frame_vjp and
 *    mdir::frame_hvp (python/mdir/torch.py:139-364).
 *
 *    The parameters theta are the per-atom charges and the per-type sigma
 *    and epsilon of the Lennard-Jones pairs, combined by the
 *    Lorentz-Berthelot rules (the type-pair tables of the description must
 *    follow those rules from their diagonal wherever epsilon is not zero).
 *    The 1-4 pair parameters are given per pair and are not tunable.
 *    Arrays of frames are frame-major: positions [count][n][3].  Frame
 *    arrays (positions, forces, cotangents, reference forces, v, out) may be
 *    host or device pointers; the others are host pointers.
 * ------------------------------------------------------------------- */

/* A frame evaluator for up to max_frames frames per call; NULL on error
   (mdrtHipLastError).  The description's positions, velocities, waters and
   skins are not used. */
void *mdrtHipFramesCreate(const mdrtHipSystemDesc *desc, int32_t max_frames);
void mdrtHipFramesDestroy(void *frames);

/* Sets (any non-NULL of) the charges [n] and the per-type sigma [ntypes]
   and epsilon [ntypes]; the loaded frames must be evaluated again. */
int mdrtHipFramesSetParameters(void *frames, const double *charges, const double *sigma,
                               const double *epsilon);
int mdrtHipFramesGetParameters(void *frames, double *charges, double *sigma,
                               double *epsilon);

/* Energies [count][MDRT_NUM_ENERGIES] (kinetic 0), forces [count][n][3]
   and the gradient in the strain [count][3][3] of `count` frames (any
   output may be NULL).  The strain gradient is S_ab = dU/d(eps_ab) for
   x -> (1 + eps) x, h -> h (1 + eps)^T, row a, column b, kJ/mol: minus the
   virial tensor, with the host's 1/V terms (net-charge background,
   dispersion correction) on the diagonal (the reference's
   gradients["strain"], docs/python-frames.md D252).  Loads the frames for
   mdrtHipFramesVjp and mdrtHipFramesHvp. */
int mdrtHipFramesEvaluate(void *frames, int32_t count, const double *positions,
                          double *energies, double *forces, double *strain);

/* The product with the parameters, summed over the loaded frames,
     p = sum_n [ g_n dU_n/dtheta + sum_i G_ni . d2U_n/(dx_i dtheta) ],
   for energy cotangents g [count] (NULL: zeros) and cotangents G
   [count][n][3] on the position gradients dU/dx (NULL: zeros; the
   cotangent on the forces with the opposite sign): the reference's
   FrameEnergies.vjp(cotangent, positions=G) (_frames.py:327).  Outputs
   d_charges [n], d_sigma [ntypes], d_epsilon [ntypes] (any may be NULL;
   d_epsilon of a type with epsilon 0 is reported as 0). */
int mdrtHipFramesVjp(void *frames, int32_t count, const double *energy_cotangent,
                     const double *position_cotangent, double *d_charges, double *d_sigma,
                     double *d_epsilon);

/* The same product in individual type-pair table entries (a, b) (the
   reference's "per pair of types" sigma and epsilon tunables, NBFIX
   entries; sigma_ab = sigma_ba and eps_ab = eps_ba are one parameter each)
   instead of the per-type parameters: 1 to 4 pairs type_pairs [npairs][2],
   outputs d_sigma [npairs], d_epsilon [npairs]. */
int mdrtHipFramesVjpPairs(void *frames, int32_t count, const double *energy_cotangent,
                          const double *position_cotangent, int32_t npairs,
                          const int32_t *type_pairs, double *d_sigma, double *d_epsilon);

/* The PME plane transform of the frame evaluator (as mdrtHipSetPmeTransform). */
int mdrtHipFramesSetPmeTransform(void *frames, int32_t mode);

/* out[n] = H_n v[n], H_n the Hessian of U_n in the positions, for the
   loaded frames (the reference's frame_hvp, torch.py:341; with the
   reciprocal PME sum included). */
int mdrtHipFramesHvp(void *frames, int32_t count, const double *v, double *out);

/* Force matching in one call: evaluates the frames, the loss
   loss[n] = sum_i |F_ni - F_ref,ni|^2 and its parameter gradient (the
   product above with G = -2 (F - F_ref), g = 0).  graph != 0 captures the
   sequence in a HIP graph at the first call with this count and replays
   it afterwards. */
int mdrtHipFramesForceMatching(void *frames, int32_t count, const double *positions,
                               const double *reference_forces, int32_t graph, double *loss,
                               double *d_charges, double *d_sigma, double *d_epsilon);

/* Launch-shape trial of the frame evaluator (as mdrtHipAutotune, for the
   frame batch): times the graph-replayed mdrtHipFramesForceMatching on
   these `count` frames for each shape `what` names (as mdrtHipAutotune),
   one untimed call and `repeats` timed ones each in two passes, and keeps
   the fastest (the current shape unless another is 2 % faster).
   ms_per_call[24] as ms_per_step there. The evaluator holds no evaluated
   frames afterwards. */
int mdrtHipFramesAutotune(void *frames, int32_t what, int32_t count, const double *positions,
                          const double *reference_forces, int32_t repeats, double *ms_per_call);

/* ---------------------------------------------------------------------
 * 4. Replicas: independent systems of uneven size advanced together, each
 *    chunk of steps of all of them one replayed HIP graph (a branch per
 *    replica).  No counterpart in the reference, whose runs take one
 *    system; its device batch (lib/Compiler/Simulation.cpp:3189) batches
 *    frames of one system.
 * ------------------------------------------------------------------- */

/* A set of `count` systems (each as mdrtHipSystemCreate); NULL on error. */
void *mdrtHipReplicasCreate(const mdrtHipSystemDesc *descs, int32_t count);
void mdrtHipReplicasDestroy(void *set);

/* The system handle of replica `index`, owned by the set (not to be
   destroyed): every group-2 entry point takes it (positions, forces,
   Langevin, launch shapes, the launch-shape trial). */
void *mdrtHipReplicasSystem(void *set, int32_t index);

/* `steps` steps of dt (ps) of every replica, in chunks of `graph` steps,
   each chunk of all replicas one graph launch.  A replica whose list
   outgrew its buffers redoes its chunk alone with the grown list. */
int mdrtHipReplicasRun(void *set, int64_t steps, double dt, int32_t graph);

/* How a chunk is launched: 0 (the default) one graph holding every
   replica's branch, on one stream; 1 one graph per replica (the same
   branch), each on the replica's own stream, so that the runtime's
   hardware queues (GPU_MAX_HW_QUEUES) run replicas side by side. */
int mdrtHipReplicasSetLaunch(void *set, int32_t mode);

/* Step latency of replica `index` inside the batch: for each chunk since the
   last clear, (device clock at the end of its branch - at its start) /
   steps, ms.  Writes up to `capacity` values; *count receives how many
   there are. */
int mdrtHipReplicasLatency(void *set, int32_t index, int64_t capacity, double *ms_per_step,
                           int64_t *count);
int mdrtHipReplicasClearLatency(void *set);

/* ---------------------------------------------------------------------
 * 5. No counterpart in the reference: for callers that cannot allocate
 *    device memory (code under HIP graph capture, tight serving loops,
 *    harnesses that forbid hipMalloc).
 * ------------------------------------------------------------------- */

/* _mlir_ciface_mdrtCudaFFTForward3DF32 without a plan: the real grid
   k1 x k2 x k3 (f32, device) to its half spectrum k1 x k2 x (k3/2+1)
   (interleaved complex f32, device), forward exp(-2 pi i), unnormalised,
   on `stream` (NULL: the default stream).  The three axes' twiddle tables
   are its only workspace and live in a fixed device array of the library
   (3 x 1024 complex f32 = 24 KiB per GPU, held until the process ends);
   it calls no allocator.  Edges in [2, 1024] with factors 2, 3, 5, 7 only,
   k3 even; anything else returns -1 (mdrtHipLastError).  Calls that
   overlap on one GPU (different streams) race on the tables. */
int mdrtHipFFTForward3DF32NoAlloc(const float *real, void *complex, int64_t k1, int64_t k2,
                                  int64_t k3, void *stream);

/* ---------------------------------------------------------------------
 * 6. Outputs (lib/Driver/Trajectory.cpp:48-206, lib/Driver/Output.cpp
 *    frame thread): DCD files as the reference writes them — the CORD
 *    record (frames, first step, period, frames x period, the time step in
 *    AKMA units as f32, the cell flag, version 24), one title line, the
 *    particle count, then per frame the cell (a, cos gamma, b, cos beta,
 *    cos alpha, c in Angstrom, f64) and x, y, z records of f32 Angstrom.
 * ------------------------------------------------------------------- */

/* A trajectory of `system` at `path`: a frame every `period` steps from
   step `first` (the step of the first frame), time step in ps.  keep < 0
   starts a new file; keep >= 0 continues a file this writer wrote with the
   same period and cell flag, keeping its first `keep` frames (a restart
   from a checkpoint that counted them).  NULL on failure
   (mdrtHipLastError). */
void *mdrtHipTrajectoryOpen(void *system, const char *path, int64_t first, int64_t period,
                            double timestep_ps, int64_t keep);
/* Stages the system's positions now (after the steps issued so far) as the
   frame of step `step`; a thread of the writer converts and writes it
   while the next steps run.  At most 3 frames are in flight; the call of a
   fourth waits for the oldest.  A copy that failed on the writer's thread
   is reported here, through mdrtStagingFail. */
int mdrtHipTrajectoryFrame(void *trajectory, int64_t step);
/* Waits until every staged frame is in the file. */
int mdrtHipTrajectoryFlush(void *trajectory);
/* Frames the file holds once the staged ones are written. */
int64_t mdrtHipTrajectoryFrames(void *trajectory);
/* Flushes, closes the file and frees the slots. */
int mdrtHipTrajectoryClose(void *trajectory);

/* A whole DCD file at once from host arrays: `frames` frames of `n`
   particles, positions in nm (frames x n x 3, f64), the orthorhombic cell
   of each frame in nm (frames x 3; NULL: no cell records), converted as the
   trajectory writer converts them (f32 Angstrom). */
int mdrtHipDcdWrite(const char *path, int64_t n, int64_t frames, int64_t first,
                    int64_t period, double timestep_ps, const double *positions,
                    const double *cells);

#ifdef __cplusplus
}
#endif
#endif
