"""GPU: the batched ensemble (one launch per kernel over all members) gives
the same bits as stepping each member alone, for members of different grid
sizes, discharges, roughness and tides, with and without sediment."""
import pytest

from hydist_amd import synth
from hydist_amd._lib import Ensemble, HydistError, Solver

from _parity import assert_same_state

pytestmark = pytest.mark.gpu


def _integrates(case, steps=25):
    g = Solver(case)
    try:
        g.init()
        g.run(0.0, steps, case.opts())
        return True
    except HydistError:
        return False
    finally:
        g.close()


def small_members(sediment, seed):
    return synth.ensemble_cases(n=3, seed=seed, side_min=64, side_max=160, sediment=sediment, accept=_integrates)


@pytest.mark.parametrize("sediment", [False, True])
def test_ensemble_matches_members(sediment):
    cases = small_members(sediment, seed=11 if sediment else 7)
    op = cases[0].opts()
    alone = [Solver(c) for c in cases]
    batch = [Solver(c) for c in cases]
    for g in alone + batch:
        g.init()
    for g in alone:
        g.run(0.0, 25, op)
    ens = Ensemble(batch)
    t = ens.run(0.0, 25, op)
    assert t == 25 * cases[0].dT
    for k, (a, b) in enumerate(zip(alone, batch)):
        assert_same_state(a, b, f"member {k} ({cases[k].name})")


def test_ensemble_counted_wet_cells():
    cases = small_members(False, seed=5)
    op = cases[0].opts()
    ms = [Solver(c) for c in cases]
    for g in ms:
        g.init()
    ens = Ensemble(ms)
    ens.reset_wet_counter()
    ens.run(0.0, 10, op, counted=True)
    total = ens.wet_cell_steps()
    assert total > 0
    # the counter holds the wet cells of each step's end state; the last
    # step's end state is the current mask
    assert total >= sum(m.wet_count() for m in ms)


def test_ensemble_rejects_uninitialised_member():
    c = small_members(False, seed=3)[0]
    g = Solver(c)
    with pytest.raises(HydistError) as e:
        Ensemble([g])
    assert e.value.code == -6


def test_tuned_launch_cached_and_state_unchanged(tmp_path, monkeypatch):
    """hydist_amd.tuning: a case is timed once and then set from the cache;
    tuning an ensemble leaves its members' states as they were, and the
    tuned ensemble still gives each member's own bits."""
    import numpy as np
    from hydist_amd import tuning
    monkeypatch.setenv("HYDIST_TUNE_CACHE", str(tmp_path / "launch.json"))
    cases = small_members(True, seed=11)
    op = cases[0].opts()
    g = Solver(cases[0])
    g.init()
    first = tuning.tune(g, 0.0, op, steps=4)
    again = tuning.tune(g, 0.0, op, steps=4)
    assert not first["cached"] and again["cached"]
    assert (again["rb_mult"], again["bw"]) == (first["rb_mult"], first["bw"])
    assert first["rb_mult"] in tuning.RB_CHOICES and first["bw"] in tuning.BW_CHOICES
    g.close()
    alone = [Solver(c) for c in cases]
    batch = [Solver(c) for c in cases]
    for s in alone + batch:
        s.init()
    before = [s.get("z").copy() for s in batch]
    info = tuning.tune_ensemble(batch, 0.0, op, steps=4)
    assert len(info["trials"]) == len(tuning.RB_CHOICES) * len(tuning.BW_CHOICES)
    for s, z in zip(batch, before):
        assert np.array_equal(s.get("z"), z)
    assert tuning.tune_ensemble(batch, 0.0, op, steps=4)["cached"]
    for s in alone:
        s.run(0.0, 20, op)
    ens = Ensemble(batch)
    ens.run(0.0, 20, op)
    for k, (a, b) in enumerate(zip(alone, batch)):
        assert_same_state(a, b, f"member {k} after tuning")
