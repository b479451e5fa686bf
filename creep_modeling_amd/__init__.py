"""MI355X-native LS-EVPFFT per-increment fixed-point loop.

The package holds:

* ``inputs``  -- writers for the solver's input files (microstructure ``.dat``,
  ``cupl2.sx``, ``cuel.sx``, ``fft.in``) and the seeded synthetic RVE generator.
* ``parse``   -- readers for the same files (the solver's input contract),
  producing a :class:`~creep_modeling_amd.parse.Problem`.
* ``lib``     -- ctypes binding of ``libcreep_hip.so`` (the C-ABI declared in
  ``include/creep_hip.h``).
* ``driver``  -- the increment / fixed-point loop of ``LS-EVPFFT.f90`` driving
  the HIP kernels through the C-ABI, plus the ``results.csv`` /
  ``diagnostics.csv`` writers.

The Fortran host executable (``host/``) is the drop-in for
``lsevpfft/LS-EVPFFT``; the Python driver mirrors it for tests and benchmarks.
"""

__all__ = ["inputs", "parse", "lib", "driver"]
