"""The oracle: CPU restatements of SIMP's hot path, used only as a checker.

TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package.  The product
(``simp_amd``) never imports, links or executes anything here.

* :mod:`oracle.capi` -- ctypes binding of ``liboracle_simp.so``
  (``simp_oracle.c``, FP64 restatement of the reference routines).
* :mod:`oracle.refrun` -- driver of the compiled reference itself
  (``_ref/simp_ref_harness``, built from the reference sources by
  ``build_ref.sh``), used to write golden vectors and as the CPU baseline.
"""
