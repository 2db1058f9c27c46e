"""GPU: the flow surrogate end to end on a small river: the device-resident
training set from plain FP64 ADI runs, training with the library's own
kernels, what the trained network gains over the method without it (the
persistence prediction), and the accelerated bed-change driver with its FP64
check and fallback."""
import numpy as np
import pytest

from hydist_amd import cnn, synth
from hydist_amd import surrogate as S
from hydist_amd._lib import Solver
from hydist_amd.case import BED_EVERY_STEP

pytestmark = pytest.mark.gpu

SCALE = 0.1
K = 20


def small_net(n_samples):
    c = S.river_variant(0, scale=SCALE)
    net = cnn.Network(c.N + 3, c.M + 3, batch=2, layers=8, seed=3)
    net.set_scales(**S.SCALES)
    return net


def test_network_beats_persistence():
    net = small_net(24)
    n = S.build_dataset(net, variants=2, pairs=12, K=K, scale=SCALE, log=lambda s: None)
    hist, val, va = S.train(net, n, epochs=12, batch=2, lr=1e-3, val_frac=0.15, log=lambda s: None)
    pers = S.persistence_loss(net, va)
    assert np.isfinite(hist).all() and np.isfinite(val)
    assert hist[-1] < hist[0]
    assert val < pers, (val, pers)


def test_surrogate_driver_checks_and_falls_back():
    case = synth.config(3, scale=SCALE, hours=6.0)
    case.bed_mode = BED_EVERY_STEP
    net = small_net(2)
    op = case.opts()
    g = Solver(case)
    g.init()
    tb = 2.0 * 3600.0
    g.run(0.0, int(round(tb / case.dT)), op)
    # an impossible bound (a negative fraction of cells): every check fails,
    # the last verified state comes back and plain ADI takes over from it
    rep = S.surrogate_run(case, tb, tb + 30 * K * case.dT, g, net, K, check_every=2, z_frac=-1.0,
                          fallback_macro=3)
    assert rep.checks >= 1 and rep.fallbacks == rep.checks
    assert rep.adi_steps >= 3 * K * rep.fallbacks
    assert abs(rep.sim_hours - 30 * K * case.dT / 3600.0) < 1e-9
    # with every surrogate call rejected, the run is plain ADI: same bed as a
    # plain run over the same span
    ref = Solver(case)
    ref.init()
    ref.run(0.0, int(round(tb / case.dT)) + 30 * K, op)
    assert np.array_equal(g.get("dH"), ref.get("dH"))
    # morphological acceleration keeps the time line
    g2 = Solver(case)
    g2.init()
    g2.run(0.0, int(round(tb / case.dT)), op)
    rep2 = S.surrogate_run(case, tb, tb + 6 * K * case.dT, g2, net, K, check_every=100, morph_substeps=4)
    assert rep2.surrogate_calls == 6 and abs(rep2.sim_hours - 6 * K * case.dT / 3600.0) < 1e-9


def test_surrogate_accepted_at_real_bounds():
    """With the stated bounds (at most 0.1 % of the wet cells moved by more
    than 5 mm in the FP64 check step, wet volume within 1e-4) a network
    trained on the river's own pairs is accepted on the small river at
    K = 10, and the bed it drives stays close to the plain FP64 run's: 99 %
    of the wet cells within 15 mm and 90 % within 1 mm after an hour of bed
    change every step (the network's outputs are BF16, so not bit-exact)."""
    K10 = 10
    net = small_net(80)
    n = S.build_dataset(net, variants=2, pairs=40, K=K10, scale=SCALE, log=lambda s: None)
    S.train(net, n, epochs=20, batch=2, lr=1e-3, val_frac=0.1, log=lambda s: None)
    case = synth.config(3, scale=SCALE, hours=6.0, SL=S.SEGMENT_LIMIT)
    case.bed_mode = BED_EVERY_STEP
    op = case.opts()
    plain = Solver(case)
    plain.init()
    tb = 2.0 * 3600.0
    plain.run(0.0, int(round(tb / case.dT)), op)
    sur = plain.clone()
    S.plain_run(case, tb, tb + 3600.0, plain)
    rep = S.surrogate_run(case, tb, tb + 3600.0, sur, net, K10, check_every=2, z_bound=5e-3, z_frac=1e-3,
                          mass_bound=1e-4)
    assert rep.checks >= 20 and rep.accepted >= 0.75 * rep.checks, rep
    assert rep.surrogate_calls >= 40 and abs(rep.sim_hours - 1.0) < 1e-9
    wet = plain.get("khouot") == 0
    d = S.bed_difference(plain.get("dH"), sur.get("dH"), wet)
    assert d["p99_m"] <= 0.015 and d["frac_within_1mm"] >= 0.9, d
    assert np.isfinite(sur.get("z")[wet]).all()


def test_perfect_flow_scheme():
    """The scheme with the true flow in place of the network: the flow does
    not depend on the bed (dH is never fed back into the depths), so u, v, z
    equal a plain run's bit for bit after every macro step, and the bed run
    on the flow frozen over each macro step stays close to the plain run's."""
    case = synth.config(3, scale=SCALE, hours=6.0, SL=S.SEGMENT_LIMIT)
    case.bed_mode = BED_EVERY_STEP
    op = case.opts()
    plain = Solver(case)
    plain.init()
    tb = 2.0 * 3600.0
    plain.run(0.0, int(round(tb / case.dT)), op)
    pf = plain.clone()
    S.plain_run(case, tb, tb + 3600.0, plain)
    rep = S.perfect_flow_run(case, tb, tb + 3600.0, pf, 10)
    assert rep.macro_steps == 72 and abs(rep.sim_hours - 1.0) < 1e-9
    for f in ("u", "v", "z", "khouot"):
        assert np.array_equal(plain.get(f), pf.get(f)), f
    wet = plain.get("khouot") == 0
    d = S.bed_difference(plain.get("dH"), pf.get("dH"), wet)
    assert np.isfinite(pf.get("dH")[wet]).all()
    assert d["p99_m"] <= 0.005, d      # measured 1.7 mm
