"""ctypes binding of libmdrt_hip.so (include/mdrt_hip.h).

The library is built in-tree by mdir_amd.build; importing this module does
not build it.  Every entry point fails loudly when the library is missing:
there is no CPU fallback of the force path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmdrt_hip.so")

E_LJ, E_COUL_DIRECT, E_COUL_EXCL, E_COUL_RECIP, E_COUL_SELF, E_DISP_CORR, \
    E_LJ14, E_COUL14, E_KINETIC = range(9)
NUM_ENERGIES = 14
TERM_PAIRS, TERM_SPECIAL, TERM_RECIP, TERM_BONDED = 1, 2, 4, 8
ENERGY_NAMES = ["lj", "coul_direct", "coul_excl", "coul_recip", "coul_self",
                "disp_corr", "lj14", "coul14", "kinetic", "bonds", "urey_bradley",
                "angles", "dihedrals", "impropers"]
KINETIC = 8

D, I32, I64 = C.c_double, C.c_int32, C.c_int64
PD, PI32, PI64 = C.POINTER(D), C.POINTER(I32), C.POINTER(I64)


class RunState(C.Structure):
    """mdrtHipRunState: a run's state beyond positions and velocities."""
    _fields_ = [("step_count", I64), ("box", D * 3), ("barostat_dv", D),
                ("barostat_counts", I64 * 4), ("list_kept", I32), ("reserved", I32)]


class SystemDesc(C.Structure):
    """mdrtHipSystemDesc."""
    _fields_ = [("n", I64), ("positions", PD), ("velocities", PD),
                ("masses", PD), ("charges", PD), ("types", PI32),
                ("ntypes", I32), ("sigma", PD), ("epsilon", PD),
                ("box", D * 3), ("cutoff", D), ("switch_from", D),
                ("lj_modifier", I32), ("coulomb_shift", I32),
                ("dispersion", I32), ("beta", D), ("pme_grid", I32 * 3),
                ("pme_order", I32), ("nexclusions", I64), ("exclusions", PI32),
                ("npairs14", I64), ("pairs14", PI32), ("pairs14_sigma", PD),
                ("pairs14_eps", PD), ("pairs14_qq", PD), ("nwaters", I64),
                ("water_first", PI32), ("d_oh", D), ("d_hh", D), ("skin", D),
                ("prune_skin", D), ("ntuples", I64), ("tuple_atoms", PI32),
                ("tuple_kind", PI32), ("tuple_param", PD), ("nclusters", I64),
                ("cluster_first", PI32), ("cluster_size", PI32), ("cluster_center", PI32),
                ("cluster_length", PD)]


class Buffer1(C.Structure):
    """mdrtBuffer1: the one-dimensional memref descriptor."""
    _fields_ = [("allocated", C.c_void_p), ("aligned", C.c_void_p),
                ("offset", I64), ("size", I64), ("stride", I64)]


class LibraryMissing(RuntimeError):
    pass


_lib = None

# (name, restype, argtypes) of every entry point of include/mdrt_hip.h
SIGNATURES = [
    ("mgpuSetDefaultDevice", None, [I32]),
    ("mgpuModuleLoad", C.c_void_p, [C.c_void_p, C.c_size_t]),
    ("mgpuModuleLoadJIT", C.c_void_p, [C.c_void_p, C.c_int, C.c_size_t]),
    ("mgpuModuleUnload", None, [C.c_void_p]),
    ("mgpuModuleGetFunction", C.c_void_p, [C.c_void_p, C.c_char_p]),
    ("mgpuLaunchKernel", None, [C.c_void_p, C.c_ssize_t, C.c_ssize_t, C.c_ssize_t,
                                C.c_ssize_t, C.c_ssize_t, C.c_ssize_t, I32,
                                C.c_void_p, C.POINTER(C.c_void_p),
                                C.POINTER(C.c_void_p), C.c_size_t]),
    ("mgpuLaunchKernelCooperative", None, [C.c_void_p] + [C.c_ssize_t] * 9 +
     [I32, C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    ("mdrtCooperativeGridSize", C.c_ssize_t, [C.c_ssize_t]),
    ("mgpuStreamCreate", C.c_void_p, []),
    ("mgpuStreamDestroy", None, [C.c_void_p]),
    ("mgpuStreamSynchronize", None, [C.c_void_p]),
    ("mgpuStreamWaitEvent", None, [C.c_void_p, C.c_void_p]),
    ("mgpuEventCreate", C.c_void_p, []),
    ("mgpuEventDestroy", None, [C.c_void_p]),
    ("mgpuEventSynchronize", None, [C.c_void_p]),
    ("mgpuEventRecord", None, [C.c_void_p, C.c_void_p]),
    ("mgpuMemAlloc", C.c_void_p, [C.c_uint64, C.c_void_p, C.c_bool]),
    ("mgpuMemFree", None, [C.c_void_p, C.c_void_p]),
    ("mgpuMemcpy", None, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    ("mdrtCudaMemcpyScoped", None, [C.c_size_t, C.c_size_t, C.c_uint64]),
    ("mdrtCudaWaitScoped", None, []),
    ("mgpuMemset32", None, [C.c_void_p, C.c_uint, C.c_size_t, C.c_void_p]),
    ("mgpuMemset16", None, [C.c_void_p, C.c_ushort, C.c_size_t, C.c_void_p]),
    ("mdrtSideBegin", None, []),
    ("mdrtSideEnd", None, []),
    ("mdrtSide2Begin", None, []),
    ("mdrtSide2End", None, []),
    ("mdrtSideJoin", None, []),
    ("mdrtUseStream", C.c_void_p, [C.c_void_p]),
    ("mdrtRestoreStream", None, [C.c_void_p]),
    ("_mlir_ciface_mdrtFlagMemory", None, [C.POINTER(Buffer1), I64]),
    ("mdrtFlagStart", I64, [C.c_void_p]),
    ("mdrtFlagFinish", I32, [I64]),
    ("mdrtFlagTake", I32, [C.c_void_p]),
    ("_mlir_ciface_mdrtHostResults", None, [C.POINTER(Buffer1), I64, I64]),
    ("mdrtCountHostRead", None, [C.c_size_t, C.c_int, C.c_int]),
    ("mdrtCudaSetStopHandler", None, [C.c_void_p]),
    ("mdrtSetFailureHook", None, [C.c_void_p]),
    ("mdrtDeviceTrim", C.c_uint64, []),
    ("mdrtDeviceMemoryCounts", None, [PI64]),
    ("mdrtDeviceAllocateKept", C.c_void_p, [C.c_uint64]),
    ("mdrtDeviceFreeKept", None, [C.c_void_p]),
    ("mdrtDeviceCopyWithin", None, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("mdrtDeviceCopyToHost", None, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("mdrtStagingCreate", C.c_void_p, []),
    ("mdrtStagingCopy", C.c_void_p, [C.c_void_p, I32, C.c_void_p, C.c_uint64]),
    ("mdrtStagingEnd", None, [C.c_void_p]),
    ("mdrtStagingWait", I32, [C.c_void_p, C.POINTER(C.c_char_p)]),
    ("mdrtStagingFail", None, [I32, C.c_char_p]),
    ("mdrtStagingDestroy", None, [C.c_void_p]),
    ("_mlir_ciface_mdrtCudaFFTForward3D", None,
     [C.POINTER(Buffer1), C.POINTER(Buffer1), I64, I64, I64]),
    ("_mlir_ciface_mdrtCudaFFTBackward3D", None,
     [C.POINTER(Buffer1), C.POINTER(Buffer1), I64, I64, I64]),
    ("_mlir_ciface_mdrtCudaFFTForward3DF32", None,
     [C.POINTER(Buffer1), C.POINTER(Buffer1), I64, I64, I64]),
    ("_mlir_ciface_mdrtCudaFFTBackward3DF32", None,
     [C.POINTER(Buffer1), C.POINTER(Buffer1), I64, I64, I64]),
    ("_mlir_ciface_mdrtCudaFFTGaussianBackwardF32", None,
     [C.POINTER(Buffer1), C.POINTER(Buffer1), C.POINTER(Buffer1), I64, I64, I64]),
    ("mdrtRetainFFTPlan", C.c_int, [I64, I64, I64, C.c_int, C.c_int, C.c_void_p]),
    ("mdrtReleaseFFTPlan", None, [C.c_int]),
    ("mdrtHipSystemCreate", C.c_void_p, [C.POINTER(SystemDesc)]),
    ("mdrtHipSystemDestroy", None, [C.c_void_p]),
    ("mdrtHipLastError", C.c_char_p, []),
    ("mdrtHipSetPositions", C.c_int, [C.c_void_p, PD]),
    ("mdrtHipSetVelocities", C.c_int, [C.c_void_p, PD]),
    ("mdrtHipGetPositions", C.c_int, [C.c_void_p, PD]),
    ("mdrtHipGetVelocities", C.c_int, [C.c_void_p, PD]),
    ("mdrtHipComputeForces", C.c_int, [C.c_void_p, PD, PD]),
    ("mdrtHipComputeForceTerms", C.c_int, [C.c_void_p, I32, PD, PD]),
    ("mdrtHipSetPmeTransform", C.c_int, [C.c_void_p, I32]),
    ("mdrtHipAutotune", C.c_int, [C.c_void_p, I32, I32, I32, D, PD]),
    ("mdrtHipSetLaunchShape", C.c_int, [C.c_void_p, I32, I32, I32]),
    ("mdrtHipGetLaunchShape", C.c_int, [C.c_void_p, PI32]),
    ("mdrtHipFramesSetPmeTransform", C.c_int, [C.c_void_p, I32]),
    ("mdrtHipSetLangevin", C.c_int, [C.c_void_p, D, D, C.c_uint64]),
    ("mdrtHipSetStepCount", C.c_int, [C.c_void_p, I64]),
    ("mdrtHipListPairs", C.c_int, [C.c_void_p, D, I32, I64, PI32, PI64]),
    ("mdrtHipFFTForward3DF32NoAlloc", C.c_int, [C.c_void_p, C.c_void_p, I64, I64, I64,
                                                C.c_void_p]),
    ("mdrtHipReplicasCreate", C.c_void_p, [C.POINTER(SystemDesc), I32]),
    ("mdrtHipReplicasDestroy", None, [C.c_void_p]),
    ("mdrtHipReplicasSystem", C.c_void_p, [C.c_void_p, I32]),
    ("mdrtHipReplicasRun", C.c_int, [C.c_void_p, I64, D, I32]),
    ("mdrtHipReplicasSetLaunch", C.c_int, [C.c_void_p, I32]),
    ("mdrtHipReplicasLatency", C.c_int, [C.c_void_p, I32, I64, PD, PI64]),
    ("mdrtHipReplicasClearLatency", C.c_int, [C.c_void_p]),
    ("mdrtHipGetBox", C.c_int, [C.c_void_p, PD]),
    ("mdrtHipScaleBox", C.c_int, [C.c_void_p, PD]),
    ("mdrtHipBarostatMove", C.c_int, [C.c_void_p, D, D, C.c_uint64, PI32, PD]),
    ("mdrtHipRunNPT", C.c_int, [C.c_void_p, I64, D, I32, I32, D, D, C.c_uint64, PD]),
    ("mdrtHipBarostatStats", C.c_int, [C.c_void_p, PI64, PD]),
    ("mdrtHipGetStepCount", C.c_int, [C.c_void_p, PI64]),
    ("mdrtHipSaveState", C.c_int, [C.c_void_p, PD, PD, PD, PD, PD, C.POINTER(RunState)]),
    ("mdrtHipRestoreState", C.c_int, [C.c_void_p, PD, PD, PD, PD, C.POINTER(RunState)]),
    ("mdrtHipTrajectoryOpen", C.c_void_p, [C.c_void_p, C.c_char_p, I64, I64, D, I64]),
    ("mdrtHipTrajectoryFrame", C.c_int, [C.c_void_p, I64]),
    ("mdrtHipTrajectoryFlush", C.c_int, [C.c_void_p]),
    ("mdrtHipTrajectoryFrames", I64, [C.c_void_p]),
    ("mdrtHipTrajectoryClose", C.c_int, [C.c_void_p]),
    ("mdrtHipDcdWrite", C.c_int, [C.c_char_p, I64, I64, I64, I64, D, PD, PD]),
    ("mdrtHipRun", C.c_int, [C.c_void_p, I64, D, I32, PD]),
    ("mdrtHipPrepareRun", C.c_int, [C.c_void_p, I64, D, I32]),
    ("mdrtHipCheckList", C.c_int, [C.c_void_p, PD, PD, PI64]),
    ("mdrtHipRunMeasured", C.c_int, [C.c_void_p, I64, D, I32, PD]),
    ("mdrtHipMinimize", C.c_int, [C.c_void_p, I64, D, PD]),
    ("mdrtHipNeighborPairs", C.c_int, [C.c_void_p, D, I32, I64, PI32, PI64]),
    ("mdrtHipStats", C.c_int, [C.c_void_p, PI64]),
    ("mdrtHipProfileStep", C.c_int, [C.c_void_p, I32, PD]),
    ("mdrtHipFramesCreate", C.c_void_p, [C.POINTER(SystemDesc), I32]),
    ("mdrtHipFramesDestroy", None, [C.c_void_p]),
    ("mdrtHipFramesSetParameters", C.c_int, [C.c_void_p, PD, PD, PD]),
    ("mdrtHipFramesGetParameters", C.c_int, [C.c_void_p, PD, PD, PD]),
    ("mdrtHipFramesEvaluate", C.c_int, [C.c_void_p, I32, C.c_void_p, PD, C.c_void_p, PD]),
    ("mdrtHipFramesVjp", C.c_int, [C.c_void_p, I32, PD, C.c_void_p, PD, PD, PD]),
    ("mdrtHipFramesVjpPairs", C.c_int, [C.c_void_p, I32, PD, C.c_void_p, I32, PI32, PD, PD]),
    ("mdrtHipFramesHvp", C.c_int, [C.c_void_p, I32, C.c_void_p, C.c_void_p]),
    ("mdrtHipFramesAutotune", C.c_int, [C.c_void_p, I32, I32, C.c_void_p, C.c_void_p, I32,
                                        PD]),
    ("mdrtHipFramesForceMatching", C.c_int, [C.c_void_p, I32, C.c_void_p, C.c_void_p, I32,
                                             PD, PD, PD, PD]),
]


def lib():
    """The loaded library; raises LibraryMissing when it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is missing: build it with `python -m mdir_amd.build` "
                "(the force path has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != 0:
        raise RuntimeError("mdrt_hip: " + lib().mdrtHipLastError().decode())


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def system_desc(s, skin: float = 0.1, prune_skin: float = 0.0):
    """The mdrtHipSystemDesc of a System and the arrays it points into
    (keep them alive until the create call returns)."""
    keep = []

    def k(a, dtype):
        a = np.ascontiguousarray(a, dtype)
        keep.append(a)
        return a

    x = k(s.positions, np.float64)
    v = k(s.velocities if s.velocities is not None else np.zeros_like(x), np.float64)
    d = SystemDesc()
    d.n = s.n
    d.positions = ptr(x, D)
    d.velocities = ptr(v, D)
    d.masses = ptr(k(s.masses, np.float64), D)
    d.charges = ptr(k(s.charges, np.float64), D)
    d.types = ptr(k(s.types, np.int32), I32)
    d.ntypes = s.ntypes
    d.sigma = ptr(k(s.sigma, np.float64), D)
    d.epsilon = ptr(k(s.epsilon, np.float64), D)
    for a in range(3):
        d.box[a] = float(s.box[a])
        d.pme_grid[a] = int(s.pme_grid[a])
    d.cutoff = s.cutoff
    d.switch_from = s.switch_from
    d.lj_modifier = s.lj_modifier
    d.coulomb_shift = int(s.coulomb_shift)
    d.dispersion = int(s.dispersion)
    d.beta = s.beta
    d.pme_order = s.pme_order
    ex = k(s.exclusions.reshape(-1, 2), np.int32)
    d.nexclusions = ex.shape[0]
    d.exclusions = ptr(ex, I32)
    p14 = k(s.pairs14.reshape(-1, 2), np.int32)
    d.npairs14 = p14.shape[0]
    d.pairs14 = ptr(p14, I32)
    d.pairs14_sigma = ptr(k(s.pairs14_sigma, np.float64), D)
    d.pairs14_eps = ptr(k(s.pairs14_eps, np.float64), D)
    d.pairs14_qq = ptr(k(s.pairs14_qq, np.float64), D)
    w = k(s.waters, np.int32)
    d.nwaters = w.shape[0]
    d.water_first = ptr(w, I32)
    d.d_oh, d.d_hh = s.water_geometry
    d.skin = skin
    d.prune_skin = prune_skin
    ta, tk, tp = s.tuples()
    d.ntuples = tk.shape[0]
    d.tuple_atoms = ptr(k(ta, np.int32), I32)
    d.tuple_kind = ptr(k(tk, np.int32), I32)
    d.tuple_param = ptr(k(tp, np.float64), D)
    cl, length = s.clusters()
    d.nclusters = cl.shape[0]
    d.cluster_first = ptr(k(cl[:, 0], np.int32), I32)
    d.cluster_size = ptr(k(cl[:, 1], np.int32), I32)
    d.cluster_center = ptr(k(cl[:, 2], np.int32), I32)
    d.cluster_length = ptr(k(length, np.float64), D)
    return d, keep
