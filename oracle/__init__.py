"""FP64 CPU oracle for the pixel-space covariance path (test infrastructure only).

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import it.
It restates, in NumPy/SciPy and plain C, the reference algorithms of
ggalloni/CosmoForge 1.3.0 for:

* HEALPix pixel centres (``healpix``; healpy is absent, Gorski et al. 2005),
* a minimal FITS reader for the reference's fixture files (``fitsio``),
* the spin-0 / spin-2 / spin-0x2 Legendre recurrences, rotation angles and
  the signal / binned-derivative matrices (``kernels`` + ``csrc``),
* the Jacobi-scaled Cholesky inverse and log-determinant (``linalg``),
* Fisher, QML spectra / noise bias and PICSLike chi2 + logdet (``pipeline``).

Each function cites the reference file:line it follows.  Pins (golden
vectors, closed forms, invariants) live in ``tests/test_oracle_*.py``.
"""
