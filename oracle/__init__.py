"""CPU oracle for the FRI-CCSD hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or as the timed
CPU baseline.  The product package (``fri-cc_amd/fricc``) never imports it.

Contents
--------
``liboracle``   ctypes bindings to ``libfricc_oracle.so`` (compression in the
                canonical and faithful modes, the 11 reference contraction loops,
                the fma-chain ERI builder).
``rccsd``       NumPy FP64 restatement of ``fricc/friccsd.py:update_amps`` and
                the PySCF formulas it calls (``imd.cc_Foo`` ... ``energy``,
                ``init_amps``).
``spinorb``     An independent spin-orbital CCSD (Stanton/Gauss form) used to
                pin ``rccsd`` without PySCF.
"""
