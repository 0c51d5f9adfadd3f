"""Serving loop on the device: continuous batching over concurrent
streams, HBM-footprint admission, and results that do not depend on the
mix (each query alone and among others gives the same masses)."""
import numpy as np
import pytest

from effmass_amd import serving

pytestmark = pytest.mark.gpu


def _queries():
    Q = serving.MassQuery
    return [Q(qid=0, natom=2, ecut_ry=30.0, nb=16, kfracs=[np.array([0.1, 0.2, -0.3])], seed=5),
            Q(qid=1, natom=8, ecut_ry=30.0, nb=24, kfracs=[np.array([0.0, 0.1, 0.2]), np.array([0.3, 0.0, 0.0])],
              seed=6),
            Q(qid=2, natom=2, ecut_ry=40.0, nb=20, kfracs=[np.array([-0.2, 0.1, 0.05])], seed=7),
            Q(qid=3, natom=8, ecut_ry=40.0, nb=32, kfracs=[np.array([0.15, -0.1, 0.3])], seed=8)]


def test_same_result_alone_and_in_the_mix():
    """Each query alone (its k-points one after the other) and served in the
    mix (a 2-k-point query runs as two parts on two slots) gives the same
    masses to 1e-10 relative and energies to 1e-12 Ha."""
    qs = _queries()
    alone = {q.qid: serving.run_query(q) for q in (qs[0], qs[1], qs[3])}
    ans, stats = serving.serve(qs, slots=4, realtime=False)
    print(stats, flush=True)
    assert stats["errors"] == 0
    assert stats["max_concurrency"] >= 2
    for qid, a in alone.items():
        mix = ans[qid]
        assert len(mix.inv_mass) == len(a.inv_mass) and mix.npw == a.npw
        for x, y in zip(a.inv_mass, mix.inv_mass):
            assert np.abs(np.asarray(x) - np.asarray(y)).max() <= 1e-10 * np.abs(np.asarray(x)).max()
        assert np.allclose(a.energies, mix.energies, rtol=0, atol=1e-12)


def test_admission_by_hbm_footprint():
    qs = _queries()
    need = max(serving.estimate_bytes(q) for q in qs)
    ans, stats = serving.serve(qs, slots=4, realtime=False, hbm_budget=need)
    print(stats, flush=True)
    assert stats["errors"] == 0
    assert stats["peak_admitted_bytes"] <= need      # never more admitted than the budget
    assert stats["max_concurrency"] < 4              # the budget held some queries back
    big = serving.MassQuery(qid=9, natom=64, ecut_ry=50.0, nb=256, kfracs=[np.zeros(3)], seed=1)
    ans2, stats2 = serving.serve([qs[0], big], slots=2, realtime=False, hbm_budget=need)
    assert ans2[9].error and "budget" in ans2[9].error
    assert ans2[0].error is None


def test_footprint_estimate_covers_a_large_cell_with_few_bands():
    """The largest projector share per band: 64 atoms at 50 Ry with 16
    bands (projectors and their k-gradients at k and at the shifted k of
    the effective-mass step) runs within its own estimate."""
    q = serving.MassQuery(qid=0, natom=64, ecut_ry=50.0, nb=16, kfracs=[np.array([0.1, -0.2, 0.3])], seed=3)
    a = serving.run_query(q)                  # budget = estimate_bytes(q)
    assert a.error is None and len(a.inv_mass) == 1


def test_query_mass_is_converged():
    """The answer is a property of the query, not of the solver's start: the
    mixed-precision m* equals the one from a tightly converged FP64 solve
    (other start vectors, nb + 8 of nb + 16 rows converged to 1e-10) within 1e-4
    relative, and the band energy within 1e-6 Ry."""
    import torch

    from effmass_amd import effmass, eigensolver
    from effmass_amd.hamiltonian import Hamiltonian

    q = serving.MassQuery(qid=0, natom=8, ecut_ry=30.0, nb=24, kfracs=[np.array([0.12, -0.05, 0.2])], seed=11)
    ans = serving.run_query(q)
    assert ans.error is None
    wl = serving.build_workload(q)
    kp = wl.kpts[0]
    H = Hamiltonian(wl.cell, device=0)
    try:
        H.load_kpoint(0, kp)
        kin = torch.from_numpy(kp.kin).to("cuda:0")
        X0 = eigensolver.random_start(q.nb + 16, kp.npw, kin, seed=1234, device="cuda:0")
        op = H.at_k(0)
        res = eigensolver.chfsi(lambda V, HV: op(V, HV, prec="fp64", overwrite=True), X0, kin, nconv=q.nb + 8,
                                tol=1e-10, maxiter=400, gram=H.overlap, degree=20)
        assert res.converged == q.nb + 8
        band = q.nb // 2
        ref = effmass.effective_mass(H, 0, kp, wl.cell.A, res.X[:q.nb].contiguous(), res.E[:q.nb], [band],
                                     prec="fp64").inv_mass[0]
    finally:
        H.close()
    got = np.asarray(ans.inv_mass[0])
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"m* rel diff {rel:.2e}, dE {abs(ans.energies[0] - float(res.E[band])):.2e} Ha", flush=True)
    assert rel <= 1e-4
    assert abs(ans.energies[0] - float(res.E[band])) <= 0.5e-6      # Ha (1e-6 Ry)
