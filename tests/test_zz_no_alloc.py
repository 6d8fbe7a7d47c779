"""smolamdInjectDisplacementsNoAlloc, the injected step that allocates no
device memory (deviates staged in a fixed device array), against
smolamdInjectDisplacements, the entry point it varies: two calls in a row on
twin simulations, the second with a smaller input, bit-exact states and
reaction events; the size limit and the set-up requirement are reported as
errors."""
import numpy as np
import pytest

import simspec as ss

pytestmark = pytest.mark.gpu


def _twins(n, seed):
    spec = ss.c1_spec(n=n, seed=seed)
    return spec, ss.build_device(spec), ss.build_device(spec)


def _ndiff(sim):
    return sum(len(sim.molecules(ll)["ident"]) for ll in range(sim.nlists()))


def test_no_alloc_step_equals_allocating_step():
    rng = np.random.default_rng(5)
    # first call: 6000 molecules; second call: another simulation of 1500
    for n, seed in ((3000, 41), (750, 42)):
        spec, a, b = _twins(n, seed)
        for _ in range(2):
            g = rng.standard_normal((_ndiff(a), 3))
            a.inject_step(g)
            b.inject_step_noalloc(g)
            assert ss.canonical_pairs(a.reaction_log()) == ss.canonical_pairs(b.reaction_log())
            ss.compare_states(ss.device_state(a), ss.device_state(b))


def test_no_alloc_limits_are_errors():
    import smoldyn_amd as sa
    spec = ss.c1_spec(n=200, seed=43)
    sim = ss.build_device(spec, update=False)
    with pytest.raises(sa.SmoldynError) as e:
        sim.inject_step_noalloc(np.zeros((400, 3)))
    assert "smolUpdateSim first" in str(e.value)
    sim.update()
    big = 262144 + 1
    with pytest.raises(sa.SmoldynError) as e:
        sim.inject_step_noalloc(np.zeros((big, 3)))
    assert "allocation-free limit" in str(e.value)
    # the simulation still steps afterwards
    sim.inject_step_noalloc(np.zeros((400, 3)))
