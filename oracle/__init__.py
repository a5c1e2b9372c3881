"""CPU oracle of the AOSpectrum Band path -- TEST INFRASTRUCTURE ONLY.

This package restates the reference's algorithm on the host (NumPy/SciPy) so
that the HIP implementation in ``aospectrum_amd`` can be checked against it.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the reported CPU
baseline -- never as part of the product path.

Modules and what pins each of them (beyond the reference citations inside):

* ``oracle.bloch``  -- Bloch assembly (src/aospectrum/sparse.py:63-245); pinned
  against brute-force dense Bloch sums and the invariants H(k)^H = H(k),
  H(k + G) = H(k), H(-k) = conj H(k) for real blocks.
* ``oracle.ciss``   -- contour geometry, probes, moments, compression,
  Rayleigh-Ritz, ghost filter, selection, residual
  (src/aospectrum/ciss.py:81-391) with a SciPy SuperLU shifted solver; pinned
  end to end by the FP64 dense generalized eigensolver (exact in-window counts,
  eigenvalues within 1e-4 eV) and by the scalar quadrature identity of the
  contour filter.
* ``oracle.dense``  -- FP64 dense ``eigh(H(k), S(k))`` reference spectrum and
  Sylvester-inertia counts; pinned by scipy.linalg's LAPACK drivers.
"""
