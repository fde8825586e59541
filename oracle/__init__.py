"""CPU oracle: an FP64 numpy restatement of amaterasu's H.psi / Davidson / density path.

THIS PACKAGE IS TEST INFRASTRUCTURE. It exists only to check the HIP product path.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and there only as the checker (or the timed CPU
baseline) -- never as the thing measured or shipped. Nothing under
``amaterasu_amd/`` imports it.

Every function cites the reference file:line whose algorithm it restates
(paths relative to the amaterasu source tree, package version 0.1.0a2).
The reference itself cannot be imported anywhere in this build (its hard
dependencies ase/spglib/h5py are absent), so the oracle is pinned by the
reference's own closed-form tests re-expressed in ``tests/test_oracle_*.py``
(empty lattice, FFT normalisation, Parseval, Y_lm addition theorem, projector
completeness, Davidson vs dense ``eigh``, ...). Functions that no such check
pins are listed as "parity unpinned" in their docstrings.
"""
