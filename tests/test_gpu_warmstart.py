"""Learned warm start (extension; the reference always starts cold,
src/lp/first_order/pdlp.cpp:186-191) on the GPU: the BF16 MLP is trained on
GPU PDLP solutions of a perturbed LP family and its predictions start the
test instances.

Bars:
* every warm-started instance reaches 1e-6 (the cold fallback catches the
  ones that do not within the iteration cap), and its returned pair re-scored
  by the oracle's FP64 residuals is <= 1.5e-6;
* what the trained network gains: the warm start's mean iteration count is
  at least 10 % below the cold start's (measured on this family: 8,062 ->
  5,652, -29.9 %; the gain is printed);
* with an iteration cap too small for any instance, every instance falls back
  to the cold start and still converges.
"""
import numpy as np
import pytest

import mcpdlp as mp
import oracle as orc
import warmstart

pytestmark = pytest.mark.gpu

KW = dict(tolerance=1e-6, backend=mp.BACKEND_GPU, require_gpu=1, restart_scheme=mp.SCHEME_HALPERN,
          precision=mp.PRECISION_MIXED, max_iterations=400000)


@pytest.fixture(scope="module")
def trained():
    ntr, nval, nte = 80, 16, 16
    base, fam = mp.generate_family(dict(rows=500, cols=1250, nnz_min=6, nnz_max=6),
                                   seeds=[0x4D4350444C50 + 9000 + k for k in range(ntr + nval + nte)], sigma=0.1)
    train, val, test = fam[:ntr], fam[ntr:ntr + nval], fam[ntr + nval:]
    print("warm start family", base.rows, base.cols, base.nnz, len(fam), flush=True)
    sols = mp.solve_batch(train, **KW)
    assert all(r["status"] == "optimal" for r in sols)
    ws = warmstart.WarmStart(base)
    fit = ws.fit(train, [r["primal"] for r in sols], [r["dual"] for r in sols],
                 omegas=[r["primal_weight"] for r in sols], iterations=[r["iterations"] for r in sols])
    tune = ws.tune_omega(val, **KW)
    print("fit", fit, "tune", tune, flush=True)
    return ws, test


def test_warm_start_converges_and_does_not_lose(trained):
    ws, test = trained
    cold = mp.solve_batch(test, **KW)
    warm, fallbacks = ws.solve_batch(test, **KW)
    for lp, r in zip(test, warm):
        assert r["status"] == "optimal"
        assert orc.kkt(lp, r["primal"], r["dual"])["score"] <= 1.5e-6
    ci = np.mean([r["iterations"] for r in cold])
    wi = np.mean([r["iterations"] for r in warm])
    print(f"iterations to 1e-6: cold mean {ci:.0f}, warm mean {wi:.0f} ({100 * (wi / ci - 1):+.1f} %), "
          f"{fallbacks} cold fallbacks", flush=True)
    assert wi <= 0.9 * ci


def test_cold_fallback_catches_every_unconverged_instance(trained):
    ws, test = trained
    cap = ws.cap
    try:
        ws.cap = 1  # one iteration: no check runs, so no warm start can converge
        warm, fallbacks = ws.solve_batch(test[:4], **KW)
    finally:
        ws.cap = cap
    assert fallbacks == 4
    for lp, r in zip(test[:4], warm):
        assert r["status"] == "optimal" and r.get("warm_fallback")
        assert orc.kkt(lp, r["primal"], r["dual"])["score"] <= 1.5e-6
