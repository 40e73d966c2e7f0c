"""CPU tests of the host side: the library loads and exports every symbol
include/gsqg.h and include/gsqg_closure.h declare (no compute calls), and the
host tables (wavenumbers, masks, flag packing, ring list, keys, initial
conditions) match the oracle."""

import os
import re

import numpy as np
import pytest

from oracle import ics as oics
from oracle import reference_sqg as R
from oracle import rng as orng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    syms = set()
    for h in ("gsqg.h", "gsqg_closure.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        syms |= set(re.findall(r"^\s*(?:int|const char\*)\s+(gsqg_\w+)\s*\(", txt, re.M))
    return sorted(syms)


def test_library_exports_every_header_symbol():
    from pygsquig_amd import _lib

    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.EXPORTED)
    assert lib.gsqg_version() >= 100


def test_missing_library_fails_loudly(tmp_path):
    from pygsquig_amd import _lib

    with pytest.raises(_lib.GsqgLibraryError):
        _lib.load(str(tmp_path / "nope.so"))


@pytest.mark.parametrize("N,L", [(16, 2 * np.pi), (256, 2 * np.pi), (64, 3.0), (1024, 2 * np.pi)])
def test_wavenumbers_and_masks_match_oracle(N, L):
    from pygsquig_amd.grid import wavenumbers_1d
    from pygsquig_amd.plan import half_plane_masks

    k = wavenumbers_1d(N, L)
    np.testing.assert_array_equal(k, R.wavenumbers_1d(N, L))
    g = R.make_grid(N, L)
    kf, dk = (N / 8.0, 2.0)
    d, ls, rg = half_plane_masks(N, L, damping_kf=kf, ring=(kf, dk))
    np.testing.assert_array_equal(d, g.dealias_mask[:, : N // 2 + 1])
    np.testing.assert_array_equal(ls, R.large_scale_mask(g, kf)[:, : N // 2 + 1])
    np.testing.assert_array_equal(rg, R.RingForcing(kf, dk).forcing_mask(g)[:, : N // 2 + 1])


def test_flag_packing():
    from pygsquig_amd.plan import pack_flags

    N = 16
    rs = np.random.default_rng(0)
    d, ls, rg = (rs.random((N, N // 2 + 1)) < 0.5 for _ in range(3))
    f = pack_flags(d, ls, rg)
    assert f.shape == (N // 2, N) and f.dtype == np.uint8
    for j in range(1, N // 2):
        for i in range(N):
            assert f[j, i] == (d[i, j] | (ls[i, j] << 1) | (rg[i, j] << 2))
    for i in range(N):
        assert f[0, i] & 15 == (d[i, 0] | (ls[i, 0] << 1) | (rg[i, 0] << 2))
        assert f[0, i] >> 4 == (d[i, N // 2] | (ls[i, N // 2] << 1) | (rg[i, N // 2] << 2))


def test_keys_match_oracle():
    from pygsquig_amd import rng

    for s, n in [(0, 0), (42, 7), (2**63 + 5, 123456)]:
        assert rng.step_key(s, n) == orng.step_key(s, n)
    assert rng.key_to_u64(np.array([3, 4], np.uint32)) == orng.key_to_u64(np.array([3, 4], np.uint32))


def test_initial_conditions_match_oracle():
    from pygsquig_amd import ics

    np.testing.assert_array_equal(ics.kolmogorov_53(64, 3, alpha=0.5), oics.kolmogorov_53(64, 3, alpha=0.5))
    np.testing.assert_array_equal(ics.sin4(32), oics.sin4(32))
    np.testing.assert_array_equal(ics.gaussian_blob(32, (1.0, 2.0), 0.4),
                                  oics.gaussian_blob(32, (1.0, 2.0), 0.4))
    th = ics.kolmogorov_53(128, 9)
    assert abs(np.sqrt(np.mean(th**2)) - 1.0) < 1e-12


def test_solver_validation_messages():
    import pygsquig_amd as pg

    class FakeGrid:
        N, L = 16, 2 * np.pi

    with pytest.raises(ValueError, match=r"alpha must be in \[-2, 2\], got 3"):
        pg.gSQGSolver(FakeGrid(), 3)
    with pytest.raises(ValueError, match="p must be 2, 4, or 8, got 3"):
        pg.gSQGSolver(FakeGrid(), 1.0, p=3)
    with pytest.raises(ValueError, match="p must be 2, 4, or 8"):
        pg.CombinedDamping(p=5)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message (no GPU
    is touched before the check)."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
