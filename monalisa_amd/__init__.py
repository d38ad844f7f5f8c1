"""monalisa_amd -- MI355X (gfx950) implementation of Monalisa's non-Cartesian
reconstruction hot path.

The compute runs in ``libmona.so`` (hand-written HIP kernels behind a C ABI,
declared in ``include/mona.h``).  This package is the host-side mirror of the
reference's MATLAB interface for that path:

* :mod:`monalisa_amd.sparse`    -- ``bmSparseMat`` (jump-encoded CSR) and
  ``bmTraj2SparseMat``;
* :mod:`monalisa_amd.nufft`     -- ``bmShanna`` / ``bmNakatsha`` / ``bmNasha``
  operators and the ``bmSensa`` / ``bmSteva`` / ``bmTevaMorphosia_chain``
  solvers on device-resident plans;
* :mod:`monalisa_amd.synth`     -- seeded trajectories, volume elements, coil
  maps and phantoms for the benchmark workloads;
* :mod:`monalisa_amd.serve`     -- the continuous-batching reconstruction
  server.

Importing the package never falls back to a CPU implementation: without the
built library (or without a GPU for device calls) the calls raise.
"""
from ._lib import lib, MonaError, check  # noqa: F401

__all__ = ["lib", "MonaError", "check"]
