"""ctypes front of the FP64 oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package mdir_amd.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
REF_FFT = os.path.join(HERE, "_ref", "libmdrt_fft_ref.so")
REF_DRAWS = os.path.join(HERE, "_ref", "libref_draws.so")
REF_HOST = os.path.join(HERE, "_ref", "libmdrt_host_ref.so")

E_LJ, E_COUL_DIRECT, E_COUL_EXCL, E_COUL_RECIP, E_COUL_SELF, E_DISP_CORR, \
    E_LJ14, E_COUL14, E_BOND, E_UREY_BRADLEY, E_ANGLE, E_DIHEDRAL, E_IMPROPER = range(13)
NTERMS = 13
TERM_NAMES = ["lj", "coul_direct", "coul_excl", "coul_recip", "coul_self",
              "disp_corr", "lj14", "coul14", "bonds", "urey_bradley", "angles",
              "dihedrals", "impropers"]

_lib = None


def build(ref: bool = True) -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        _lib = C.CDLL(LIB)
        _declare(_lib)
    return _lib


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


D, I32, I64 = C.c_double, C.c_int32, C.c_int64
PD, PI32, PI64 = C.POINTER(D), C.POINTER(I32), C.POINTER(I64)


class OrcSystem(C.Structure):
    _fields_ = [("n", I64), ("q", PD), ("type", PI32), ("ntypes", I32),
                ("sigma", PD), ("epsilon", PD), ("mass", PD), ("box", D * 3),
                ("cutoff", D), ("switch_from", D), ("lj_modifier", I32),
                ("coulomb_shift", I32), ("dispersion", I32), ("beta", D),
                ("coulomb", D), ("excl_start", PI32), ("excl_atoms", PI32),
                ("nexcl_pairs", I64), ("excl_pairs", PI32), ("K", I64 * 3),
                ("order", I32), ("nwaters", I64), ("water_first", PI32),
                ("m_o", D), ("m_h", D), ("d_oh", D), ("d_hh", D),
                ("npairs14", I64), ("pairs14", PI32), ("p14_sigma", PD), ("p14_eps", PD),
                ("p14_qq", PD), ("ntuples", I64), ("tuple_atoms", PI32),
                ("tuple_kind", PI32), ("tuple_param", PD), ("nclusters", I64),
                ("cluster_first", PI32), ("cluster_size", PI32), ("cluster_center", PI32),
                ("cluster_length", PD)]


def _declare(L):
    L.orc_cell_count.restype = I64
    L.orc_cell_count.argtypes = [D, D]
    L.orc_cell_range.restype = I64
    L.orc_cell_range.argtypes = [D, I64, D]
    L.orc_cell_width.restype = D
    L.orc_cell_width.argtypes = [I64, PD, D, D]
    L.orc_neighbors_matrix.restype = I64
    L.orc_neighbors_matrix.argtypes = [I64, PD, PD, D, D, PI32, PI32, I64]
    L.orc_exclusion_filter.argtypes = [I64, PI32, PI32, I64, PI32, PI32]
    L.orc_nonbonded.restype = C.c_int
    L.orc_nonbonded.argtypes = [I64, PD, PD, PI32, I32, PD, PD, PD, D, D, I32,
                                I32, D, D, PI32, PI32, I32, PD, PD, PD]
    L.orc_excluded.argtypes = [I64, PI32, PD, PD, PD, D, D, PD, PD]
    L.orc_pairs14.argtypes = [I64, PI32, PD, PD, PD, PD, PD, D, D, I32, PD,
                              PD, PD]
    L.orc_ewald_beta.restype = D
    L.orc_ewald_beta.argtypes = [D, D]
    L.orc_bonded.restype = None
    L.orc_bonded.argtypes = [I64, PI32, PI32, PD, PD, PD, PD, PD]
    L.orc_shake_positions.restype = None
    L.orc_shake_positions.argtypes = [C.POINTER(OrcSystem), PD, PD]
    L.orc_rattle_velocities.restype = None
    L.orc_rattle_velocities.argtypes = [C.POINTER(OrcSystem), PD, PD]
    L.orc_smooth_size.restype = I64
    L.orc_smooth_size.argtypes = [I64]
    L.orc_self_energy.restype = D
    L.orc_self_energy.argtypes = [I64, PD, D, D, PD]
    L.orc_dispersion_energy.restype = D
    L.orc_dispersion_energy.argtypes = [I64, PI32, I32, PD, PD, I64, PI32, PD, D]
    L.orc_bspline.argtypes = [D, C.c_int, PD, PD]
    L.orc_spline_moduli.argtypes = [I64, C.c_int, PD]
    L.orc_pme_spread.argtypes = [I64, PD, PD, PD, PI64, C.c_int, PD]
    L.orc_pme_influence.argtypes = [PD, PI64, C.c_int, D, D, PD]
    L.orc_pme_gather.argtypes = [I64, PD, PD, PD, PI64, C.c_int, PD, PD]
    L.orc_pme.restype = D
    L.orc_pme.argtypes = [I64, PD, PD, PD, PI64, C.c_int, D, D, PD, PD]
    L.orc_fft_line.argtypes = [I64, PD, C.c_int]
    L.orc_dft_line.argtypes = [I64, PD, PD, C.c_int]
    L.orc_fft3d_r2c.argtypes = [PI64, PD, PD]
    L.orc_fft3d_c2r.argtypes = [PI64, PD, PD]
    L.orc_ewald_reciprocal.restype = D
    L.orc_ewald_reciprocal.argtypes = [I64, PD, PD, PD, D, D, PD]
    L.orc_settle_positions.argtypes = [I64, PI32, D, D, D, D, PD, PD]
    L.orc_settle_velocities.argtypes = [I64, PI32, D, D, PD, PD]
    L.orc_total.restype = C.c_int
    L.orc_total.argtypes = [C.POINTER(OrcSystem), PD, PD, PD, I32]
    L.orc_nve.restype = C.c_int
    L.orc_nve.argtypes = [C.POINTER(OrcSystem), PD, PD, D, I64, PD, I32]
    L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32)]
    L.orc_langevin_normals.argtypes = [C.c_uint64, I64, C.c_uint32, PD]
    L.orc_langevin.restype = C.c_int
    L.orc_langevin.argtypes = [C.POINTER(OrcSystem), PD, PD, D, I64, I64, C.c_uint64, D, D,
                               PD, I32]
    L.orc_md_create.restype = C.c_void_p
    L.orc_md_create.argtypes = [C.POINTER(OrcSystem), D]
    L.orc_md_destroy.argtypes = [C.c_void_p]
    L.orc_md_run.restype = C.c_int
    L.orc_md_run.argtypes = [C.c_void_p, PD, PD, D, I64]
    L.orc_md_builds.restype = I64
    L.orc_md_builds.argtypes = [C.c_void_p]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_threads.restype = C.c_int
    L.orc_threads.argtypes = []
    L.orc_splitmix64.restype = C.c_uint64
    L.orc_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
    L.orc_xoshiro256ss.restype = C.c_uint64
    L.orc_xoshiro256ss.argtypes = [C.POINTER(C.c_uint64)]
    L.orc_initial_normals.argtypes = [C.c_uint64, I64, PD]
    L.orc_initial_velocities.argtypes = [C.c_uint64, I64, PD, D, PD, I64, PI32,
                                         I64, PD]


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def neighbors_matrix(x, box, reach, cell_width=None, row_width=None):
    """Rows of the reference's matrix build; returns (counts, index, largest)."""
    L = lib()
    x, box = f64(x), f64(box)
    n = x.shape[0]
    if cell_width is None:
        cell_width = L.orc_cell_width(n, _p(box, D), reach, 0.0)
    width = row_width or 64
    while True:
        counts = np.zeros(n, np.int32)
        index = np.full((n, width), -1, np.int32)
        largest = L.orc_neighbors_matrix(n, _p(x, D), _p(box, D), reach,
                                         cell_width, _p(counts, I32),
                                         _p(index, I32), width)
        if row_width or largest <= width:
            return counts, index, largest
        width = int(largest * 1.25) + 1


def exclusion_filter(counts, index, excl_start, excl_atoms):
    L = lib()
    counts, index = i32(counts).copy(), i32(index).copy()
    L.orc_exclusion_filter(counts.shape[0], _p(counts, I32), _p(index, I32),
                           index.shape[1], _p(i32(excl_start), I32),
                           _p(i32(excl_atoms), I32))
    return counts, index


class Oracle:
    """The oracle bound to one system (mdir_amd.systems.System)."""

    def __init__(self, system):
        self.s = system
        self._keep = []
        s = system
        self.q = f64(s.charges)
        self.types = i32(s.types)
        self.sig = f64(s.sigma)
        self.eps = f64(s.epsilon)
        self.mass = f64(s.masses)
        self.excl_start, self.excl_atoms = s.exclusion_csr()
        self.excl_pairs = i32(s.exclusions.reshape(-1, 2))
        self.waters = i32(s.waters)
        st = OrcSystem()
        st.n = s.n
        st.q = _p(self.q, D)
        st.type = _p(self.types, I32)
        st.ntypes = s.ntypes
        st.sigma = _p(self.sig, D)
        st.epsilon = _p(self.eps, D)
        st.mass = _p(self.mass, D)
        for k in range(3):
            st.box[k] = float(s.box[k])
            st.K[k] = int(s.pme_grid[k])
        st.cutoff = s.cutoff
        st.switch_from = s.switch_from
        st.lj_modifier = s.lj_modifier
        st.coulomb_shift = int(s.coulomb_shift)
        st.dispersion = int(s.dispersion)
        st.beta = s.beta
        from mdir_amd.systems import COULOMB, TIP3P
        st.coulomb = COULOMB
        st.excl_start = _p(self.excl_start, I32)
        st.excl_atoms = _p(self.excl_atoms, I32)
        st.nexcl_pairs = self.excl_pairs.shape[0]
        st.excl_pairs = _p(self.excl_pairs, I32)
        st.order = s.pme_order
        st.nwaters = len(s.waters)
        st.water_first = _p(self.waters, I32)
        if len(s.waters):
            st.m_o = float(s.masses[s.waters[0]])
            st.m_h = float(s.masses[s.waters[0] + 1])
        else:
            st.m_o, st.m_h = TIP3P["m_o"], TIP3P["m_h"]
        st.d_oh, st.d_hh = s.water_geometry
        # 1-4 pairs, bonded tuples and constraint clusters
        self.p14 = i32(s.pairs14.reshape(-1, 2))
        self.p14_s, self.p14_e, self.p14_q = (f64(s.pairs14_sigma), f64(s.pairs14_eps),
                                              f64(s.pairs14_qq))
        st.npairs14 = self.p14.shape[0]
        st.pairs14 = _p(self.p14, I32)
        st.p14_sigma, st.p14_eps, st.p14_qq = (_p(self.p14_s, D), _p(self.p14_e, D),
                                               _p(self.p14_q, D))
        ta, tk, tp = s.tuples()
        self.t_atoms, self.t_kind, self.t_param = i32(ta), i32(tk), f64(tp)
        st.ntuples = self.t_kind.shape[0]
        st.tuple_atoms = _p(self.t_atoms, I32)
        st.tuple_kind = _p(self.t_kind, I32)
        st.tuple_param = _p(self.t_param, D)
        cl, length = s.clusters()
        self.cl_first = i32(cl[:, 0])
        self.cl_size = i32(cl[:, 1])
        self.cl_center = i32(cl[:, 2])
        self.cl_length = f64(length)
        st.nclusters = cl.shape[0]
        st.cluster_first = _p(self.cl_first, I32)
        st.cluster_size = _p(self.cl_size, I32)
        st.cluster_center = _p(self.cl_center, I32)
        st.cluster_length = _p(self.cl_length, D)
        self.st = st

    def total(self, x=None, mixed=False):
        """Forces (n,3) and the energy terms (ORC_NTERMS,)."""
        x = f64(self.s.positions if x is None else x)
        F = np.zeros_like(x)
        e = np.zeros(NTERMS)
        lib().orc_total(C.byref(self.st), _p(x, D), _p(F, D), _p(e, D),
                        int(mixed))
        return F, e

    def bonded(self, x=None):
        """Forces (n, 3) and energy terms (NTERMS,) of the bonded tuples
        alone."""
        x = f64(self.s.positions if x is None else x)
        F = np.zeros_like(x)
        e = np.zeros(NTERMS)
        lib().orc_bonded(self.t_kind.shape[0], _p(self.t_atoms, I32), _p(self.t_kind, I32),
                         _p(self.t_param, D), _p(x, D), _p(f64(self.s.box), D), _p(F, D),
                         _p(e, D))
        return F, e

    def shake_positions(self, x_old, x_new):
        out = f64(x_new).copy()
        lib().orc_shake_positions(C.byref(self.st), _p(f64(x_old), D), _p(out, D))
        return out

    def rattle_velocities(self, x, v):
        out = f64(v).copy()
        lib().orc_rattle_velocities(C.byref(self.st), _p(f64(x), D), _p(out, D))
        return out

    def nonbonded(self, x=None, virial=False):
        s = self.s
        x = f64(s.positions if x is None else x)
        F = np.zeros_like(x)
        e = np.zeros(NTERMS)
        w = np.zeros(9)
        from mdir_amd.systems import COULOMB
        lib().orc_nonbonded(s.n, _p(x, D), _p(self.q, D), _p(self.types, I32),
                            s.ntypes, _p(self.sig, D), _p(self.eps, D),
                            _p(f64(s.box), D), s.cutoff, s.switch_from,
                            s.lj_modifier, int(s.coulomb_shift), s.beta,
                            COULOMB, _p(self.excl_start, I32),
                            _p(self.excl_atoms, I32), 0, _p(F, D), _p(e, D),
                            _p(w, D) if virial else None)
        return (F, e, w.reshape(3, 3)) if virial else (F, e)

    def pme(self, x=None, virial=False):
        s = self.s
        x = f64(s.positions if x is None else x)
        F = np.zeros_like(x)
        w = np.zeros(6)
        K = np.array(s.pme_grid, np.int64)
        from mdir_amd.systems import COULOMB
        e = lib().orc_pme(s.n, _p(x, D), _p(self.q, D), _p(f64(s.box), D),
                          _p(K, I64), s.pme_order, s.beta, COULOMB, _p(F, D),
                          _p(w, D))
        return (F, e, w) if virial else (F, e)

    def ewald_reciprocal(self, x=None, forces=True):
        s = self.s
        x = f64(s.positions if x is None else x)
        F = np.zeros_like(x) if forces else None
        from mdir_amd.systems import COULOMB
        e = lib().orc_ewald_reciprocal(s.n, _p(x, D), _p(self.q, D),
                                       _p(f64(s.box), D), s.beta, COULOMB,
                                       _p(F, D) if forces else None)
        return (F, e) if forces else e

    def settle_positions(self, x_old, x_new):
        s = self.s
        out = f64(x_new).copy()
        lib().orc_settle_positions(len(s.waters), _p(self.waters, I32),
                                   self.st.m_o, self.st.m_h, self.st.d_oh,
                                   self.st.d_hh, _p(f64(x_old), D), _p(out, D))
        return out

    def settle_velocities(self, x, v):
        s = self.s
        out = f64(v).copy()
        lib().orc_settle_velocities(len(s.waters), _p(self.waters, I32),
                                    self.st.m_o, self.st.m_h, _p(f64(x), D),
                                    _p(out, D))
        return out

    def nve(self, x, v, dt, steps, mixed=False):
        x = f64(x).copy()
        v = f64(v).copy()
        trace = np.zeros((steps, 2))
        lib().orc_nve(C.byref(self.st), _p(x, D), _p(v, D), dt, steps,
                      _p(trace, D), int(mixed))
        return x, v, trace

    def langevin(self, x, v, dt, steps, seed, friction, temperature, first_step=0,
                 mixed=False):
        """g-BAOAB Langevin steps (friction in 1/ps, temperature in K) with
        the Philox noise of (seed, step, particle)."""
        from mdir_amd.systems import BOLTZMANN
        x = f64(x).copy()
        v = f64(v).copy()
        trace = np.zeros((steps, 2))
        lib().orc_langevin(C.byref(self.st), _p(x, D), _p(v, D), dt, steps, first_step,
                           seed, friction, BOLTZMANN * temperature, _p(trace, D), int(mixed))
        return x, v, trace

    def md(self, skin=0.15):
        """The CPU baseline's integrator (velocity Verlet with SETTLE, the
        neighbor matrix reused until an atom moved half the buffer, f32
        pair rows, fixed-point PME); call run(x, v, dt, steps) on it."""
        return MdRun(self, skin)

    def initial_velocities(self, seed, temperature):
        s = self.s
        v = np.zeros((s.n, 3))
        lib().orc_initial_velocities(seed, s.n, _p(self.mass, D), temperature,
                                     _p(f64(s.positions), D), len(s.waters),
                                     _p(self.waters, I32),
                                     s.degrees_of_freedom(), _p(v, D))
        return v


class MdRun:
    def __init__(self, oracle, skin):
        self.o = oracle
        self.h = lib().orc_md_create(C.byref(oracle.st), skin)

    def run(self, x, v, dt, steps):
        """Advances copies of x and v by `steps` steps; returns them."""
        x = f64(x).copy()
        v = f64(v).copy()
        lib().orc_md_run(self.h, _p(x, D), _p(v, D), dt, steps)
        return x, v

    @property
    def builds(self):
        return lib().orc_md_builds(self.h)

    def close(self):
        if self.h:
            lib().orc_md_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def set_threads(threads):
    lib().orc_set_threads(int(threads))


def philox4x32_10(counter, key):
    """Philox 4x32-10 of four counter words under two key words."""
    c = (C.c_uint32 * 4)(*[int(w) & 0xFFFFFFFF for w in counter])
    k = (C.c_uint32 * 2)(*[int(w) & 0xFFFFFFFF for w in key])
    out = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, out)
    return list(out)


def langevin_normals(seed, step, particle):
    r = np.zeros(3)
    lib().orc_langevin_normals(seed, step, particle, _p(r, D))
    return r


def ref_host_available():
    return os.path.exists(REF_HOST)


def ref_philox4x32(counter, key):
    """The reference's own mdrtPhilox4x32 (runtime/mdrt.c, compiled into
    oracle/_ref)."""
    L = C.CDLL(REF_HOST)
    c = (C.c_uint32 * 4)(*[int(w) & 0xFFFFFFFF for w in counter])
    k = (C.c_uint32 * 2)(*[int(w) & 0xFFFFFFFF for w in key])
    out = (C.c_uint32 * 4)()
    L.mdrtPhilox4x32(c, k, out)
    return list(out)


def initial_normals(seed, n):
    """The pre-constraint normals of the initial velocities, (n, 3)."""
    out = np.zeros((n, 3))
    lib().orc_initial_normals(seed, n, _p(out, D))
    return out


# ---- FFT helpers ----

def fft3d_r2c(real):
    real = f64(real)
    K = np.array(real.shape, np.int64)
    out = np.zeros((K[0], K[1], K[2] // 2 + 1, 2))
    lib().orc_fft3d_r2c(_p(K, I64), _p(real, D), _p(out, D))
    return out[..., 0] + 1j * out[..., 1]


def fft3d_c2r(cplx, k3):
    cplx = np.ascontiguousarray(cplx)
    K = np.array([cplx.shape[0], cplx.shape[1], k3], np.int64)
    inter = f64(np.stack([cplx.real, cplx.imag], -1))
    out = np.zeros(tuple(K))
    lib().orc_fft3d_c2r(_p(K, I64), _p(inter, D), _p(out, D))
    return out


def pme_spread(system, x=None):
    s = system
    x = f64(s.positions if x is None else x)
    K = np.array(s.pme_grid, np.int64)
    grid = np.zeros(tuple(K))
    lib().orc_pme_spread(s.n, _p(x, D), _p(f64(s.charges), D),
                         _p(f64(s.box), D), _p(K, I64), s.pme_order,
                         _p(grid, D))
    return grid


def pme_influence(system):
    from mdir_amd.systems import COULOMB
    s = system
    K = np.array(s.pme_grid, np.int64)
    bc = np.zeros((K[0], K[1], K[2] // 2 + 1))
    lib().orc_pme_influence(_p(f64(s.box), D), _p(K, I64), s.pme_order, s.beta,
                            COULOMB, _p(bc, D))
    return bc


def bspline(w, order):
    M = np.zeros(order)
    dM = np.zeros(order)
    lib().orc_bspline(w, order, _p(M, D), _p(dM, D))
    return M, dM


# ---- the reference's own host FFT (oracle/_ref) ----

class _Buffer(C.Structure):
    _fields_ = [("allocated", C.c_void_p), ("aligned", C.c_void_p),
                ("offset", I64), ("size", I64), ("stride", I64)]


def ref_fft_available():
    return os.path.exists(REF_FFT)


def ref_fft3d_r2c(real):
    """runtime/mdrt_fft.c _mlir_ciface_mdrtFFT64Forward3D (built into _ref)."""
    L = C.CDLL(REF_FFT)
    real = f64(real).copy()
    K = real.shape
    out = np.zeros((K[0], K[1], K[2] // 2 + 1, 2))
    rb = _Buffer(real.ctypes.data, real.ctypes.data, 0, real.size, 1)
    cb = _Buffer(out.ctypes.data, out.ctypes.data, 0, out.size, 1)
    L._mlir_ciface_mdrtFFT64Forward3D(C.byref(rb), C.byref(cb), I64(K[0]),
                                      I64(K[1]), I64(K[2]))
    return out[..., 0] + 1j * out[..., 1]


def ref_fft3d_c2r(cplx, k3):
    L = C.CDLL(REF_FFT)
    inter = f64(np.stack([cplx.real, cplx.imag], -1)).copy()
    K = (cplx.shape[0], cplx.shape[1], k3)
    out = np.zeros(K)
    cb = _Buffer(inter.ctypes.data, inter.ctypes.data, 0, inter.size, 1)
    rb = _Buffer(out.ctypes.data, out.ctypes.data, 0, out.size, 1)
    L._mlir_ciface_mdrtFFT64Backward3D(C.byref(cb), C.byref(rb), I64(K[0]),
                                       I64(K[1]), I64(K[2]))
    return out


# ---- the reference's own initial-velocity draws (oracle/_ref) ----

def ref_draws_available():
    return os.path.exists(REF_DRAWS)


def ref_initial_normals(seed, order):
    """lib/Driver/InitialVelocityDraws.h (built into _ref): the three normals
    of each particle of `order`, visited in that order; (n, 3) by particle."""
    L = C.CDLL(REF_DRAWS)
    L.ref_initial_normals.argtypes = [C.c_uint64, I64, PI64, PD]
    order = np.ascontiguousarray(order, np.int64)
    out = np.full((len(order), 3), np.nan)
    L.ref_initial_normals(seed, len(order), _p(order, I64), _p(out, D))
    return out


def rel_rms(a, b):
    """sqrt(sum|a-b|^2 / sum|b|^2) — test/Driver/Inputs/compare_forces_rms.py:20-22."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b * b)))
