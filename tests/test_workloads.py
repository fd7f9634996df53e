"""The seeded workload generators are reproducible and have the stated laws."""
import numpy as np

from hydro_amd import workloads as W


def test_batch_instances_pinned_and_distributed():
    runs = W.batch_instances()
    assert len(runs) == 512
    first = [(r.params.nx, round(r.params.gamma, 6), r.bx, r.by) for r in runs[:3]]
    again = [(r.params.nx, round(r.params.gamma, 6), r.bx, r.by)
             for r in W.batch_instances()[:3]]
    assert first == again
    sides = np.array([r.params.nx for r in runs])
    assert sides.min() >= 64 and sides.max() <= 1024
    assert all(r.params.nx == r.params.ny and r.params.dx == 1.0 / r.params.nx for r in runs)
    # log-uniform: the median side is near sqrt(64 * 1024) = 256
    assert 180 < np.median(sides) < 360
    g = np.array([r.params.gamma for r in runs])
    assert g.min() >= 1.3 and g.max() <= 1.67
    e = np.array([r.e * r.params.dx ** 2 for r in runs])
    assert e.min() >= 0.1 and e.max() <= 10.0
    assert all(0 <= r.bx < r.params.nx and 0 <= r.by < r.params.ny for r in runs)
    assert all(r.params.nstepmax == 100 for r in runs)


def test_serving_requests_poisson():
    runs = W.serving_requests(4000, rate=50.0)
    gaps = np.diff([r.arrival for r in runs])
    assert abs(gaps.mean() - 1 / 50.0) < 0.1 / 50.0
    sides = np.array([r.params.nx for r in runs])
    assert sides.min() >= 128 and sides.max() <= 4096
    assert all(r.params.tend == 0.1 and r.params.nstepmax == 500 for r in runs)


def test_generators_reproduce_committed_lists():
    """The batched and serving instance lists equal the committed fixture
    (tests/golden/workloads.json) bit for bit."""
    import json
    import os
    fix = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "workloads.json")))

    def rows(runs, arrivals=False):
        return [[r.params.nx, repr(r.params.gamma), r.bx, r.by, repr(r.e)]
                + ([repr(r.arrival)] if arrivals else []) for r in runs]

    assert rows(W.batch_instances(512)) == fix["batch_instances"]["rows"]
    for rate, f in fix["serving_requests"].items():
        assert rows(W.serving_requests(32, float(rate)), arrivals=True) == f["rows"]
