"""The batched fused step (nwBatchMdStep through NWReplicaSet) over replicas
of uneven size -- worm lengths, counts, boxes, drives, stiffnesses, seeds --
against every replica stepped alone through nwFusedMdStep with its own
Philox stream: positions, velocities, forces, old forces, tangents and the
generator offsets agree bit for bit, with and without the hipGraph."""
import numpy as np
import pytest

from nwamd import replicas, system
from nwamd.replicas import (NW_REPLICAS_GRAPH, NW_REPLICAS_LAUNCHES, NW_REPLICAS_ONE_PER_LAUNCH, ReplicaSet,
                            worm_physics)

pytestmark = pytest.mark.gpu

SIZES = [(60, 10), (37, 17), (200, 3), (25, 30), (90, 12), (1, 64), (300, 20)]


def small_specs():
    specs = (replicas.NWReplicaSpec * len(SIZES))()
    for r, (W, n) in enumerate(SIZES):
        s = specs[r]
        s.numFilaments, s.filamentSize = W, n
        s.boxSize = max(replicas.box_for(W * n), 12.0)
        s.driveForce = [0.1, 1.0, 3.0, 10.0, 0.5, 2.0, 1.0][r]
        s.bendingK = [1.0, 5.0, 20.0, 2.5, 12.0, 3.0, 5.0][r]
        s.seed = 1000 + 7 * r
    return specs


def state(rs):
    return [[rs.get(w, r) for w in range(5)] for r in range(rs.R)]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("form,split", [(NW_REPLICAS_GRAPH, (1, 6)), (NW_REPLICAS_GRAPH, (3, 4)),
                                        (NW_REPLICAS_LAUNCHES, (7,))])
def test_batched_step_matches_one_replica_per_launch(nwlib, form, split):
    specs, phys = small_specs(), worm_physics(gamma=2.0, kT=0.1, fdt=True)
    a, b = ReplicaSet(specs, phys), ReplicaSet(specs, phys)
    for rs in (a, b):
        rs.reset(system.NW_INIT_LATTICE_RANDOM, 0.1)
    assert all(np.array_equal(x, y) for xs, ys in zip(state(a), state(b)) for x, y in zip(xs, ys))
    for k in split:
        a.update(1e-3, k, form)
    b.update(1e-3, sum(split), NW_REPLICAS_ONE_PER_LAUNCH)
    a.sync()
    b.sync()
    names = ("pos", "vel", "tangent", "force", "forceOld")
    for r, (xs, ys) in enumerate(zip(state(a), state(b))):
        for name, x, y in zip(names, xs, ys):
            assert np.array_equal(bits(x), bits(y)), (r, SIZES[r], name)
        assert a.rng_state(r) == b.rng_state(r)
    # the worms moved and stayed finite
    p0 = a.get(0, 6)
    assert np.isfinite(p0).all()
    a.close()
    b.close()


def test_batched_replicas_keep_their_own_parameters(nwlib):
    """Each replica carries its own box, grid and drive: the mean speed along
    the tangent grows with the drive f_a (terminal speed f_a / gamma)."""
    specs = (replicas.NWReplicaSpec * 2)()
    for r, fa in enumerate((0.2, 4.0)):
        s = specs[r]
        s.numFilaments, s.filamentSize, s.boxSize = 400, 10, replicas.box_for(4000)
        s.driveForce, s.bendingK, s.seed = fa, 5.0, 99 + r
    rs = ReplicaSet(specs, worm_physics(gamma=2.0, kT=0.1, fdt=True))
    rs.reset(system.NW_INIT_LATTICE_RANDOM, 0.1)
    rs.update(1e-3, 1500, NW_REPLICAS_GRAPH)  # 3 velocity relaxation times 1 / gamma
    slow, fast = rs.observables(0)["polarSpeed"], rs.observables(1)["polarSpeed"]
    assert 0.0 < slow < 0.3 and 1.0 < fast < 2.5, (slow, fast)
    p, e = rs.params(1)
    assert e.driveForce == pytest.approx(4.0) and p.numFilaments == 400
    rs.close()
