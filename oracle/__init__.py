"""CPU oracle for the rkstiff stepping hot path -- TEST INFRASTRUCTURE ONLY.

This package is a from-scratch NumPy FP64 restatement of the reference
algorithms (rkstiff, pure Python/NumPy).  It is the checker against which the
HIP product path is compared, and nothing else:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it;
* the product package ``rkstiff_amd`` never imports it and has no CPU
  fallback -- it fails loudly when the HIP library is missing.

Modules
-------
``solvers``   SolverConfig/ETDConfig, psi functions, the seven diagonal methods
              (ETD35, ETD34, ETD4, ETD5, IF34, IF45DP, IF4), the adaptive
              accept/reject controller and both ``evolve`` drivers.
``grids``     uniform Fourier grids (reference ``grids.py:40-131``).
``problems``  the nonlinear terms (KS/KdV/Burgers rfft form, NLS fft form,
              2-D Allen-Cahn rfft2 form) and the benchmark problem builders.
``cnum``      NumPy's complex division restated (used to pin device division).
``mlp``       FP64 forward/backward of the step-size proposer MLP.

Parity status of each function is stated in its docstring: "pinned by" names
the golden vectors / closed forms that check it.
"""
