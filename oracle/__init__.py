"""FP64 CPU oracle for the Davidson H·ψ path.

TEST INFRASTRUCTURE ONLY. Nothing under ``oracle/`` is part of the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and there only as the checker (or as the timed CPU
restatement), never as the code being measured or shipped.

This package is a plain NumPy restatement of the reference algorithm
(chaeeundad/baandi, Rust). Every function cites the reference file:line it
follows. It is pinned by the reference's stored goldens (KAT-1: the synthetic
Si2 SCF of ``tests/reference/runtime/generated/reference-1``; KAT-2: the skew
one-atom CLI run of ``reference-results.json``), by closed forms (Gaussian
Hankel transforms, free-electron spectra, Ewald invariance in eta) and by dense
brute force on small cells. See ``tests/test_oracle_*.py``.

Parity status per function is recorded in each module's docstring.
"""
