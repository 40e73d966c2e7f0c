"""Generate the golden trajectories used by the GPU parity tests.

Run from the repository root: ``python tests/golden/make_golden.py``.
Every case is produced by the NumPy oracle (oracle/reference_sqg.py) from
seeded inputs; the GPU tests regenerate the same inputs and compare the final
spectra.  Stored arrays are the half plane ky in [0, N/2] of the full (N, N)
spectra (the other half is checked by the short on-the-fly tests).

Cases (the stable parity variants of the benchmark configurations):
- p2: N=256 unforced, alpha=1, nu_p=1e-16, p=8, dt=1e-3, 1000 steps,
      k^-5/3 IC (seed 5, rms 0.1).
- p3: N=256 with the forced configuration's semantics: RingForcing(kf=40,
      dk=2, eps=0.1) with per-step keys step_key(42, n), CombinedDamping(
      mu=0.1, kf=40, nu_p=1e-16, p=8), solver nu_p=0, two passive scalars
      (kappa=0, Gaussian blobs sigma=2pi/16 at (pi, pi) and (pi/2, 3pi/2)),
      k^-5/3 IC (seed 2, rms 1), dt=5e-4, 200 steps.
- ifs: integrating-factor scheme with the forced configuration's physical
      hyperviscosity at N=256 (nu = 1e-8 * (512/256)^8 = 2.56e-6, p=8; the
      explicit scheme is unstable there), forcing/damping as p3, one scalar,
      dt=5e-4, 100 steps.
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ics  # noqa: E402
from oracle import reference_sqg as R  # noqa: E402
from oracle import rng  # noqa: E402

TWO_PI = 2 * np.pi


def half(a):
    return np.ascontiguousarray(a[:, : a.shape[0] // 2 + 1])


def run_forced(N, nsteps, dt, scheme, nu_solver, damp_nu, n_scalars, seed_ic=2, seed_f=42):
    g = R.make_grid(N, TWO_PI)
    s = R.Solver(g, 1.0, nu_solver, 8, scheme=scheme)
    f = R.RingForcing(kf=40, dk=2, epsilon=0.1)
    d = R.CombinedDamping(mu=0.1, kf=40, nu_p=damp_nu, p=8)
    st = s.initialize(ics.kolmogorov_53(N, seed=seed_ic))
    centres = ics.SCALAR_CENTRES[:n_scalars]
    scal = {f"s{i}": R.fft2(ics.gaussian_blob(N, c, TWO_PI / 16)) for i, c in enumerate(centres)}
    kap = {k: 0.0 for k in scal}
    for n in range(nsteps):
        key = rng.step_key(seed_f, n)
        st = s.step(st, dt, forcing=lambda y, key=key: f(y, key=key, grid=g),
                    damping=d)
        if scal:
            scal = s.step_scalars(scal, st["theta_hat"], dt, kap)
    return st, scal


def main():
    out = {}
    t0 = time.time()
    N = 256
    g = R.make_grid(N, TWO_PI)
    s = R.Solver(g, 1.0, 1e-16, 8)
    st = s.initialize(0.1 * ics.kolmogorov_53(N, seed=5))
    for _ in range(1000):
        st = s.step(st, 1e-3)
    np.savez(os.path.join(HERE, "p2_n256.npz"), theta_hat=half(st["theta_hat"]),
             time=st["time"], step=st["step"])
    out["p2"] = time.time() - t0

    t0 = time.time()
    st, scal = run_forced(256, 200, 5e-4, "explicit", 0.0, 1e-16, 2)
    np.savez(os.path.join(HERE, "p3_n256.npz"), theta_hat=half(st["theta_hat"]),
             s0=half(scal["s0"]), s1=half(scal["s1"]),
             ring_full=st["theta_hat"][R.RingForcing(40, 2).forcing_mask(g)])
    out["p3"] = time.time() - t0

    t0 = time.time()
    st, scal = run_forced(256, 100, 5e-4, "if", 0.0, 1e-8 * (512 / 256) ** 8, 1)
    np.savez(os.path.join(HERE, "ifs_n256.npz"), theta_hat=half(st["theta_hat"]),
             s0=half(scal["s0"]))
    out["ifs"] = time.time() - t0
    print({k: round(v, 1) for k, v in out.items()})


if __name__ == "__main__":
    main()
