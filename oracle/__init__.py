"""CPU oracle for the ceviche simulation hot path.

TEST INFRASTRUCTURE ONLY.  This package is a numpy/scipy restatement of the
reference algorithms (fancompute/ceviche 0.1.3) and exists to check the HIP
implementation in ``ceviche_amd``.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker or as the timed CPU baseline -- never as a product code path.

Modules
-------
fdfd_ref        FDFD operator assembly (SC-PML, Yee differences), spsolve, field
                recovery and the closed-form eps adjoint.
                follows ceviche/derivatives.py:34-216, ceviche/fdfd.py:11-244,
                ceviche/solvers.py:50-62, ceviche/primitives.py:103-135
fdtd_ref        FDTD coefficient set-up and Yee step, float64 (or complex128 for
                complex-step differentiation).
                follows ceviche/fdtd.py:12-316, ceviche/derivatives.py:16-30,
                ceviche/utils.py:153-176, 206-214
primitives_ref  sp_mult / sp_solve / spsp_mult and their VJP/JVP rules.
                follows ceviche/primitives.py:58-259, ceviche/utils.py:13-73
numdiff         grad_num / der_num / one-sided numerical jacobian.
                follows ceviche/utils.py:108-125, ceviche/jacobians.py:54-73

How each function is pinned (the reference ships no golden field values, only
gradient self-consistency tests) is stated in each module's header and tested
in tests/test_oracle_*.py.
"""
