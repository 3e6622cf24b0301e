"""CPU oracle for the UV-parameterization hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker / reported CPU baseline.

* ``oracle.param_cpu`` restates ``engine/param.py`` (float64 numpy/scipy,
  SuperLU for the direct solves exactly as the reference's ``SPDSolver``).
* ``oracle.batch_cpu`` restates the semantics of ``engine/gpu.py``'s
  ``Batch`` (soft pin per chart, batch-global pin weight, per-chart
  flip-preventing step, Jacobi flip repair, orientation fix) with exact
  float64 solves in place of the reference's capped float32 CG.
* ``oracle.validate_cpu`` restates the verdict checks of
  ``engine/validate.py`` used as a parity criterion.

Parity pinning of each function is listed in ``tests/test_oracle.py``.
"""
