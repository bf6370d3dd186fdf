"""CPU oracle for the symmetry-reduced H.v path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import anything from this package, and only as the checker
(or, for ``cpu_baseline``, as the timed CPU restatement).  The product package
``quspin_amd`` never imports it and fails loudly when its HIP library is
missing.

Contents
--------
``qsx_oracle``  numpy/pure-Python restatement of the reference algorithms
               (basis construction, representatives, normalisations, reduced
               matrix elements, projector, CSR build, Taylor-series
               ``expm_multiply``, Lanczos, the (m, s) partition and the MLP
               reference).  Each function cites the reference file:line it
               restates.
``csrc/qsx_oracle.c``  C/OpenMP restatement of the hot loops (basis
               enumeration, per-term ``inplace_op``, CSR SpMV, Taylor loop) used
               for parity at sizes numpy cannot reach and as the CPU baseline.
``c_oracle``    ctypes loader for the C restatement.

Every oracle function is pinned by a known answer, a closed form or brute
force in ``tests/test_oracle_*.py``; the pins are listed per function in the
module headers.
"""
