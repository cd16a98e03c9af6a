"""Batched mode: independent cases advanced together (one launch per phase,
per-case Krylov state, one sweep work-group per case).  Every case must
compute exactly what it computes alone, bit for bit, in FP64 and in mixed
precision, on the banded and on the level-scheduled sweeps."""
import numpy as np
import pytest

import fvmhip as F
from fvmhip import _lib as L
from tests.cases import cavity_model, channel_model, config, same_bits

pytestmark = pytest.mark.gpu


def _cases():
    return [(cavity_model(F.cavity(16), nu=0.01), 0.0),
            (channel_model(F.channel(20, 10), nu=0.02), None),
            (cavity_model(F.cavity(24), nu=0.005), 0.0),
            (channel_model(F.channel(12, 70), nu=0.01), None)]  # 70 rows: two waves of grid rows


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
@pytest.mark.parametrize("sweep", ["band", "levels"])
def test_batch_matches_single_runs(sweep, precision, monkeypatch):
    if sweep == "levels":
        monkeypatch.setenv("FVH_SWEEP", "levels")
    else:
        monkeypatch.delenv("FVH_SWEEP", raising=False)
    n_it = 6
    cases = _cases()
    cfg = config(n_it, precision=precision)
    b = F.BatchSimulation([m for m, _ in cases], cfg, [p for _, p in cases])
    assert b.dm.case_info() == (len(cases), sweep == "band")
    nb = b.run(n_it)
    assert nb == n_it
    for k, (model, pref) in enumerate(_cases()):
        s = F.Simulation(model, cfg, pref)
        assert s.run(n_it) == nb
        for name in ("Ux", "Uy", "p"):
            assert same_bits(b.get(name, k), s.get(name)), f"case {k}: {name}"
        for a, c in zip(b.residuals(k), s.residuals()):
            assert same_bits(a[:n_it], c[:n_it]), f"case {k}: residual history"
        for fld in ("Ux", "Uy", "p"):
            assert b.stats(fld, k)["niter"] == s.stats(fld)["niter"], f"case {k}: {fld}"


def test_batch_rejects_what_it_cannot_split():
    """Per-case solves need the cases' own meshes: a plain Simulation on a
    split mesh is refused, and so is GMRES (BiCGSTAB and CG are batched)."""
    cases = _cases()[:2]
    b = F.BatchSimulation([m for m, _ in cases], config(2), [p for _, p in cases])
    with pytest.raises(F.FVHError):
        F.Simulation(b.union_model, config(2))
    cases = _cases()[:2]
    g = F.BatchSimulation([m for m, _ in cases], config(2, U_solver=F.GmresSolver), [p for _, p in cases])
    with pytest.raises(F.FVHError):
        g.run(1)


def test_batch_without_graphs_is_the_same(monkeypatch):
    """Direct enqueueing (FVH_GRAPH=0) and graph replay give the same bits."""
    out = {}
    for g in ("0", "1"):
        monkeypatch.setenv("FVH_GRAPH", g)
        cases = _cases()
        b = F.BatchSimulation([m for m, _ in cases], config(4), [p for _, p in cases])
        assert b.run(4) == 4
        out[g] = [b.get(n) for n in ("Ux", "Uy", "p")]
        assert b.dm.tuning() in (1, 2, 5)
    for a, c in zip(out["0"], out["1"]):
        assert same_bits(a, c)


def test_batch_c5_extremes_against_the_oracle():
    """The batched workload's own sizes: the smallest and the largest case
    of its seeded draw (and a channel), advanced together for 3 FP64
    iterations, every case bit for bit against the oracle run alone (fields,
    residual histories)."""
    from oracle import ref
    from tests.cases import batch_cases, rel_l2

    draw = batch_cases(64)
    sizes = [m.mesh.ncells for m, _ in draw]
    pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 1})
    n_it = 3
    cfg = config(n_it + 1)
    cases = [draw[k] for k in pick]
    print(f"cases {pick}: {[sizes[k] for k in pick]} cells", flush=True)
    b = F.BatchSimulation([m for m, _ in cases], cfg, [p for _, p in cases])
    assert b.run(n_it) == n_it
    for j, k in enumerate(pick):
        model, pref = batch_cases(64)[k]
        r = ref.RefSimulation(model, F.CaseDesc(model, cfg, pref))
        assert r.run(n_it) == n_it
        for name in ("Ux", "Uy", "p"):
            got, want = b.get(name, j), r.get(name)
            assert same_bits(got, want), f"case {k}: {name} rel {rel_l2(got, want):.3e}"
        for a, c in zip(b.residuals(j), r.residuals()):
            assert same_bits(a[:n_it], c[:n_it]), f"case {k}: residual history"


def test_batch_cases_stop_at_their_own_break():
    """Cases that reach the convergence break at different iterations: each
    keeps the fields, residual history and solve statistics of its own
    break, as its own run (which stops there) has them, bit for bit."""
    cases = [(cavity_model(F.cavity(12), nu=0.02), 0.0), (channel_model(F.channel(16, 8), nu=0.05), None),
             (cavity_model(F.cavity(20), nu=0.01), 0.0)]
    cfg = config(1000)
    b = F.BatchSimulation([m for m, _ in cases], cfg, [p for _, p in cases])
    nb = b.run(1000)
    singles = []
    for model, pref in [(cavity_model(F.cavity(12), nu=0.02), 0.0), (channel_model(F.channel(16, 8), nu=0.05), None),
                        (cavity_model(F.cavity(20), nu=0.01), 0.0)]:
        s = F.Simulation(model, config(1000), pref)
        singles.append((s, s.run(1000)))
    counts = [n for _, n in singles]
    print(f"break iterations alone: {counts}; batch ran {nb}")
    assert nb == max(counts) and len(set(counts)) > 1
    for k, (s, n) in enumerate(singles):
        for name in ("Ux", "Uy", "p"):
            assert same_bits(b.get(name, k), s.get(name)), f"case {k}: {name}"
        for a, c in zip(b.residuals(k), s.residuals()):
            assert same_bits(a, c), f"case {k}: residual history"
        for fld in ("Ux", "Uy", "p"):
            assert b.stats(fld, k)["niter"] == s.stats(fld)["niter"], f"case {k}: {fld}"
