"""Host-side parts of the serving loop: the serving query stream, cells
with 135 bohr^3 per atom, FFT grids, footprint estimates."""
import numpy as np

from effmass_amd import serving


def test_cells_and_grids():
    for n in serving.SIZES:
        A, tau = serving.query_cell(n)
        assert len(tau) == n
        assert abs(abs(np.linalg.det(A)) / n - 135.0) < 1e-9
        for ecut in serving.ECUTS:
            g = serving.grid_for(A, ecut)
            for m, a in zip(g, A):
                # the density sphere (radius 2 sqrt(ecut)) without wrap-around
                assert m >= 4 * np.sqrt(ecut) * np.linalg.norm(a) / (2 * np.pi) + 1
                assert m in serving.CT_FFT_SIZES
                k = m
                for p in (2, 3, 5):
                    while k % p == 0:
                        k //= p
                assert k == 1


def test_grid_rule_matches_the_reference_fixture():
    """The rule gives the grid of the reference's Si8 PWmat data (IN.VR on
    36^3 at ecut 25 Ry, a = 10.26 bohr) and the 90^3 / 48^3 grids of the
    64-atom (40 Ry) and 8-atom (50 Ry) cells."""
    a = 10.26314072
    assert serving.grid_for(np.diag([a] * 3), 25.0) == (36, 36, 36)
    assert serving.grid_for(np.diag([2 * a] * 3), 40.0) == (90, 90, 90)
    assert serving.grid_for(np.diag([a] * 3), 50.0) == (48, 48, 48)


def test_ct_sizes_match_the_engine():
    """CT_FFT_SIZES lists exactly the engine's compile-time plan lengths."""
    import pathlib
    import re
    src = (pathlib.Path(serving.__file__).parent / "csrc" / "fft_ct.hpp").read_text()
    plans = {int(m) for m in re.findall(r"^EFFM_PLAN2\((\d+),", src, re.M)}
    body = src[src.index("#define EFFM_CT_SIZES"):]
    body = body[:body.index("\n\n")]
    listed = {int(m) for m in re.findall(r"X\((\d+)\)", body)}
    assert set(serving.CT_FFT_SIZES) == plans == listed


def test_stream_is_seeded_poisson():
    a = serving.make_stream(200, rate=50.0, seed=5)
    b = serving.make_stream(200, rate=50.0, seed=5)
    assert [q.natom for q in a] == [q.natom for q in b]
    t = np.array([q.arrival for q in a])
    assert np.all(np.diff(t) > 0)
    assert 2.5 < t[-1] < 6.0                       # ~200 / 50 s
    assert {q.natom for q in a} == set(serving.SIZES)
    assert all(16 <= q.nb <= 256 and 1 <= len(q.kfracs) <= 4 for q in a)


def test_footprint_grows_with_size():
    small = serving.MassQuery(qid=0, natom=2, ecut_ry=30.0, nb=16, kfracs=[np.zeros(3)], seed=1)
    big = serving.MassQuery(qid=1, natom=16, ecut_ry=50.0, nb=128, kfracs=[np.zeros(3)], seed=1)
    es, eb = serving.estimate_bytes(small), serving.estimate_bytes(big)
    assert 0 < es < eb


def test_server_options_are_checked_before_any_slot_starts():
    import pytest
    with pytest.raises(ValueError, match="slots"):
        serving.MassServer(slots=0)
    with pytest.raises(ValueError, match="sync"):
        serving.MassServer(slots=1, sync="sleepy")
    # "auto" leaves the HIP runtime's choice alone (no device call)
    serving._set_sync_mode(0, "auto")


def test_rayleigh_ritz_on_host_by_default():
    import inspect
    for fn in (serving.run_query, serving._worker_main, serving.serve, serving.MassServer.__init__):
        assert inspect.signature(fn).parameters["rr_host"].default is True
