"""CPU FP64 oracle for the pseudo-spectral dynamo time step.

TEST INFRASTRUCTURE ONLY.  This package restates, in plain numpy, the
algorithms of the reference PyTorch port of MagIC (rvilim/torch-magic,
``src/magic_torch/``) for the Boussinesq Chebyshev CNAB2 step and the
LINPACK-order banded solves.  Every function cites the reference file:line it
follows.  It is imported only by ``tests/``, by ``__graft_entry__.smoke()`` and
by the ``cpu_baseline`` leg of ``bench.py``, always as the checker (or the
timed CPU baseline), never as the thing measured or shipped: the product path
in ``torch-magic_amd/magic_amd`` runs on hand-written HIP kernels and fails
loudly when its extension is missing.

Pinning (what checks each part of the oracle against something other than
itself) is listed per module in the module docstrings; see
``tests/test_oracle_*.py``.
"""
