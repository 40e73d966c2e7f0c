"""CPU oracle for the pseudo-spectral gSQG RK4 step.

TEST INFRASTRUCTURE ONLY.  Nothing under ``pygsquig_amd`` imports this package:
only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
use it, and only as the checker (or, for the baseline, as the thing timed on the
host cores beside the GPU).  It is never the thing measured for the GPU figures
and never a fallback for the product path.

Modules
-------
reference_sqg
    NumPy FP64 restatement of pygSQuiG's step with the reference's own data
    layout: full complex (N, N) spectra, c2c FFTs, ``ifft2(...).real``.  Every
    function cites the reference file:line it restates.  Also holds the
    integrating-factor RK4 scheme (new; no reference counterpart), checked by
    its exact linear decay and its convergence to the explicit scheme.
rng
    The counter-based generator used for forcing phases, shared bit-for-bit
    with the HIP kernels (integer arithmetic, pinned by known-answer vectors).
ics
    Seeded initial conditions (k^-5/3 random-phase field, Gaussian blobs).
cpu_baseline
    Timed CPU path with the reference's structure (9 c2c FFTs per RHS).

Parity status of each function is listed in the module docstrings.
"""
