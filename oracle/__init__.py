"""FP64 CPU oracle for the GIWAXSim forward-scattering path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``giwaxsim_amd`` imports this package;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker (or as the timed CPU baseline),
never as the thing measured or shipped.

The oracle is a line-by-line FP64 numpy restatement of the reference path
(tchaney97/giwaxsim), with each function citing the reference ``file:line``
it follows.  Deliberate canonicalisations (each documented where it appears):

* rotations are evaluated elementwise, left to right, without FMA, instead of
  through BLAS ``np.dot`` / ``@`` (whose last bit is machine dependent);
* frames are accumulated in phi order and detector orientations in
  (psi outer, phi, theta inner) order, instead of thread-completion order
  (which makes the reference non-deterministic and, for the detector sum,
  racy: ``tools/detector.py:298``);
* the frame binning exposes the raw ``sum`` / ``count`` grids as well as the
  mean.
"""

from .giwaxs_oracle import *  # noqa: F401,F403
