"""FP64 CPU oracle for the TRG / BTRG / HOTRG coarse-graining step.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library imports, links or
calls this package: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker (or as
the timed CPU stand-in for the Julia reference, which cannot run here).

It is a numpy restatement of the reference algorithm, function by function,
each citing the reference file:line it follows (paths relative to the
TNRKit.jl v0.5.1 source tree):

* ``ztensor``  - charge-labelled dense tensors; per-coupled-charge truncated
  SVD / Hermitian eigensolve with a global top-chi cut (MatrixAlgebraKit's
  ``svd_trunc`` / ``eigh_trunc!`` / ``truncrank`` as used at
  src/schemes/btrg.jl:61,72, src/schemes/hotrg.jl:103,110,139,146).
* ``models``   - initial tensors (src/models/ising.jl:1-67,
  src/models/potts.jl:6-81, src/models/clock.jl:1-49).
* ``schemes``  - TRG / BTRG / HOTRG ``step!``, ``finalize!``, ``run!``,
  ``free_energy`` (src/schemes/*.jl, src/utility/finalize.jl:1-14,
  src/utility/free_energy.jl:13-20, src/schemes/tnrscheme.jl:33-59).
* ``onsager``  - exact 2D Ising free energy (checks src/models/ising.jl:2).
* ``torus``    - exact partition functions of small periodic lattices by
  direct contraction, the brute-force pin for un-truncated steps.

Parity pins (what each function is checked against) are listed in the
module docstrings and exercised in tests/test_oracle_*.py.
"""
from . import ztensor, models, schemes, onsager, torus  # noqa: F401
