"""The learned part end to end on the GPU, at 128 x 128 positions (needs an MI355X).

Training data come from the batched FP64 solver; the surrogate's held-out
error is judged against those solver results; the adaptive sweep is run with
and without the surrogate and every curve is compared, at all 400 fields,
with the full FP64 solver sweep.  Bars: every curve accepted; every curve
within 2e-3 of the full sweep at every field (1e-3 is the acceptance
tolerance of the sampled checks); at least 1.5x less solver work (traced
work items) than the full sweep on these 24 materials (rough and few-free
curves are solved densely; over 200 materials the ratio is 5-6x, see the
sweep figures of bench.py); and the surrogate-guided sweep doing at most 10 %
more solver work than the plain one (the measured gain is printed; over 200
materials it is a gain, see the sweep figures of bench.py).
"""
import math

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

PPR = 128


@pytest.fixture(scope="module")
def trained():
    from elasticscattering_amd import surrogate as S

    announce("surrogate: 6000 solver runs at 128x128 positions")
    X, Y, secs, items = S.training_set(300, 10, 2, seed=4, positions_per_row=PPR)
    assert X.shape == (6000, 6) and np.isfinite(Y).all()
    ntr = 5000
    announce("surrogate: training 4 members")
    sur = S.Surrogate(members=4, seed=2)
    hist = sur.fit(X[:ntr], Y[:ntr], epochs=400, batch=256)
    yield S, sur, X[ntr:], Y[ntr:], hist
    sur.close()


def test_surrogate_trains_and_generalises(trained):
    S, sur, Xh, Yh, hist = trained
    assert hist[-1].mean() < 0.1 * hist[0].mean()
    mean, spread = sur.predict(Xh)
    ok = Yh[:, 0] > 0
    rel = np.abs(mean[ok, 0] / Yh[ok, 0] - 1)
    print(f"held-out sigma_xx relative error vs FP64 solver: median {np.median(rel):.3g}, p90 "
          f"{np.percentile(rel, 90):.3g}")
    assert np.median(rel) < 0.2
    assert spread.shape == mean.shape and np.all(spread >= 0)


def test_adaptive_sweep_against_full_solver(trained):
    S, sur, _, _, _ = trained
    B = np.logspace(-1, math.log10(60), 400)
    mats = S.draw_materials(24, seed=77)
    announce("full FP64 sweep of 24 materials x 400 fields (truth)")
    truth, _, full_items = S.solver_runs(mats, [B] * len(mats), positions_per_row=PPR)
    truth = np.stack([t[:, 0, :] for t in truth])
    reports = {}
    for name, sg in (("plain", None), ("surrogate", sur)):
        announce(f"adaptive sweep ({name})")
        est, rep = S.adaptive_sweep(mats, B, surrogate=sg, positions_per_row=PPR)
        err = S.tensor_error(est, truth)
        rep["true_max_error"] = float(err.max())
        reports[name] = rep
        print(name, rep)
        assert rep["accepted"] == len(mats)
        assert err.max() <= 2e-3
        print(f"{name}: solver work {full_items / rep['traced_items']:.2f}x less than the full sweep")
        assert 1.5 * rep["traced_items"] <= full_items
    gain = reports["plain"]["traced_items"] / reports["surrogate"]["traced_items"]
    print(f"solver work saved by the surrogate: {gain:.3f}x")
    assert reports["surrogate"]["traced_items"] <= 1.1 * reports["plain"]["traced_items"]
