"""Replicas of uneven size advanced in one graph per chunk of steps: every
replica's state after a batched Langevin run is bit for bit the state of
the same replica run alone (same seed, step numbers and launch shape: the
fixed-point force sums do not depend on what else runs on the GPU); the
forces of each replica's list after the run meet the oracle; the branch
clocks give a step latency per replica."""
import numpy as np
import pytest

from mdir_amd import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("launch", [0, 1])
def test_replicas_match_single_runs_bit_for_bit(launch):
    from mdir_amd.replicas import ReplicaSet
    from mdir_amd.simulation import Simulation
    sizes = [500, 1333, 2200, 777]
    say(f"case: {len(sizes)} replicas of {sizes} waters, 40 Langevin steps in chunks of 10, "
        f"launch mode {launch}")
    systems = []
    for k, nw in enumerate(sizes):
        s = S.tip3p_replica(nw, seed=100 + k)
        sim = Simulation(s, skin=0.15)
        sim.minimize(30, 0.01)
        s = s.copy()
        s.positions = sim.positions
        s.velocities = S.initial_velocities(s, 7 + k, 300.0)
        sim.close()
        systems.append(s)
    rs = ReplicaSet(systems)
    rs.set_launch(launch)
    for k, r in enumerate(rs):
        r.set_langevin(300.0, 1.0, 1000 + k)
    rs.run(40, 0.002, graph=10)
    for k, s in enumerate(systems):
        alone = Simulation(s, skin=0.15)
        alone.set_langevin(300.0, 1.0, 1000 + k)
        alone.run(40, 0.002, graph=10)
        xb, vb = rs[k].positions, rs[k].velocities
        xa, va = alone.positions, alone.velocities
        say(f"  replica {k} ({s.n} atoms): |dx| {np.abs(xb - xa).max():.3e}, "
            f"|dv| {np.abs(vb - va).max():.3e}, steps {rs[k].step_count}")
        assert rs[k].step_count == 40
        assert np.array_equal(xb, xa) and np.array_equal(vb, va)
        F, _, counts = rs[k].check_list()
        w = s.copy()
        w.positions = xb
        Fo, _ = O.Oracle(w).total()
        err = O.rel_rms(F, Fo)
        say(f"    list forces rel rms {err:.3e}")
        assert err < 5e-6
        lat = rs.latency(k)
        assert len(lat) == 4 and (lat > 0).all()
        alone.close()
    rs.close()
