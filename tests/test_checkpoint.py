"""Checkpoints in the reference's ASCII format (ves3d_checkpoint_*, host
code): write -> read is exact; the reader follows the reference's unpack
branches for old versions (tests/golden/checkpoint_v590_p2.txt, made by
tests/golden/make_checkpoint_fixtures.py); malformed files are reported.
The restart test (GPU) follows test/SimulationTest.cc: a run restarted from
the checkpoint of its first step ends where the uninterrupted run ends, to
the reference's 5e-7 (test/SimulationTest.cc:104)."""
import os

import numpy as np
import pytest

from ves3d_amd import checkpoint as C

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_write_read_round_trip(tmp_path):
    p, nv = 6, 3
    rng = np.random.default_rng(0)
    X = rng.standard_normal((nv, 3, 2 * p * (p + 1))) * 3.7
    prm = C.defaults(p, nv)
    prm.ts = 2.5e-4
    prm.repul_dist = 0.0314159
    prm.gravity_field[0] = 0.5
    bm = rng.uniform(0.01, 0.1, nv)
    vc = np.array([1.0, 2.0, 0.5])
    ed = rng.uniform(0, 1, nv)
    f = tmp_path / "run.chk"
    C.write(f, X, prm, bm, vc, ed)
    prm2, X2, bm2, vc2, ed2 = C.read(f)
    assert np.array_equal(X2, X)
    assert np.array_equal(bm2, bm) and np.array_equal(vc2, vc) and np.array_equal(ed2, ed)
    d1, d2 = prm.as_dict(), prm2.as_dict()
    assert d1 == d2, {k: (d1[k], d2[k]) for k in d1 if d1[k] != d2[k]}
    # the layout of Parameters::pack then EvolveSurface::pack
    text = f.read_text().split("\n")
    assert text[0] == "PARAMETERS" and "/PARAMETERS" in text and text[-2] == "/EVOLVE"
    blocks = [t for t in text if t.isupper() and t.replace("/", "").isalpha()]
    assert blocks == ["PARAMETERS", "/PARAMETERS", "EVOLVE", "VESICLEPROPS", "ARRAY", "/ARRAY",
                      "ARRAY", "/ARRAY", "ARRAY", "/ARRAY", "/VESICLEPROPS", "SURFACE",
                      "VECTORS", "SCALARS", "ARRAY", "/ARRAY", "/SCALARS", "/VECTORS",
                      "/SURFACE", "/EVOLVE"]


def test_defaults_are_the_reference_parameters():
    prm = C.defaults(16, 512).as_dict()
    assert prm["bending_modulus"] == 1e-2 and prm["repul_dist"] == 5e-2
    assert prm["rep_maxit"] == 10 and prm["rep_exponent"] == 4.0 and prm["rep_tol"] == 1e-4
    assert prm["gravity_field"] == [0.0, 0.0, -1.0]
    assert prm["scheme"] == "GloballyImplicit" and prm["rep_type"] == "PolyKReparam"
    assert prm["n_surfs"] == 512 and prm["sh_order"] == 16 and prm["diff_filter_freq"] == 10


def test_reads_the_version_590_layout():
    prm, X, bm, vc, ed = C.read(os.path.join(GOLD, "checkpoint_v590_p2.txt"))
    d = prm.as_dict()
    assert d["version"] == 590 and d["sh_order"] == 2 and d["n_surfs"] == 1
    assert d["ts"] == 1e-3 and d["rep_ts"] == 1e-3 and d["num_threads"] == 4
    assert d["scheme"] == "GloballyImplicit" and d["sht_resample_order"] == 4
    assert X.shape == (1, 3, 12)
    assert np.array_equal(X.ravel(), 0.25 * np.arange(36) - 4)
    # version <= 590 stores no vesicle properties: the parameters' values
    assert bm[0] == 2.5e-2 and vc[0] == 1.0 and ed[0] == 0.0


def test_malformed_files_are_reported(tmp_path):
    from ves3d_amd import Ves3dError
    X = np.zeros((2, 3, 2 * 4 * 5))
    f = tmp_path / "a.chk"
    C.write(f, X)
    text = f.read_text()
    bad = tmp_path / "b.chk"
    bad.write_text(text.replace("rep_maxit:", "rep_maxitt:"))
    with pytest.raises(Ves3dError) as e:
        C.read(bad)
    assert e.value.code == 51       # IOBadStream
    bad.write_text(text[: len(text) // 2])
    with pytest.raises(Ves3dError):
        C.read(bad)
    with pytest.raises(Ves3dError) as e:
        C.read(tmp_path / "missing.chk")
    assert e.value.code == 50       # IOError


@pytest.mark.gpu
def test_restart_from_checkpoint_matches_uninterrupted_run(hip, tmp_path):
    import torch
    from ves3d_amd import implicit as I
    from ves3d_amd import workloads as W
    nv, p = 2, 6
    X0 = W.suspension(W.perturbed_shapes(nv, p, amp=0.05, seed=4), 0.05, seed=4)
    kw = dict(dt=1e-3, shear_rate=1.0, precision=0, precond=1, tol=1e-8, max_iter=300)
    a = I.Stepper(nv, p, **kw)
    a.set_state(torch.from_numpy(X0).cuda())
    a.step()
    X1 = a.get_state()[0].cpu().numpy()
    f = tmp_path / "a_000001.chk"
    C.write(f, X1)
    print("restart test: checkpoint written", flush=True)
    a.step()
    X2 = a.get_state()[0].cpu().numpy()
    a.close()
    prm, Xr, bm, vc, ed = C.read(f)
    assert np.array_equal(Xr, X1)
    b = I.Stepper(nv, p, **kw)
    # area / volume targets of the uninterrupted run's start
    from oracle import ves
    a0, v0 = ves.area_volume(X0, p, 2 * p)
    b.set_state(torch.from_numpy(Xr).cuda())
    b.set_area_volume(a0, v0)
    b.step()
    Xb = b.get_state()[0].cpu().numpy()
    b.close()
    assert np.max(np.abs(Xb - X2)) < 5e-7
