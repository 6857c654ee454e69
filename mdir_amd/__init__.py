"""mdir_amd — the per-step molecular dynamics force path of mdir, native to
MI355X (gfx950): cell-sorted cluster build, nonbonded pair loop (LJ +
erfc Ewald, exclusions, 1-4 pairs), PME (spread, LDS FFT, influence,
gather) and velocity Verlet with SETTLE, as hand-written HIP kernels behind
the C ABI of libmdrt_hip.so (include/mdrt_hip.h)."""
from .systems import (System, argon_fcc, argon_config1, tip3p_box, tip3p_config2,
                      ewald_beta, smooth_size, pme_grid, COULOMB, BOLTZMANN)

__all__ = ["System", "argon_fcc", "argon_config1", "tip3p_box", "tip3p_config2",
           "ewald_beta", "smooth_size", "pme_grid", "COULOMB", "BOLTZMANN",
           "Simulation"]


def __getattr__(name):
    if name == "Simulation":
        from .simulation import Simulation
        return Simulation
    raise AttributeError(name)
