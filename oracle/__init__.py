"""CPU restatement of the BayesianGroupTesting hot path -- TEST INFRASTRUCTURE ONLY.

This package is the parity checker for the HIP provider and the timed CPU
baseline. Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it. The provider package never imports it.

Parity status of each restated function (what pins it) is listed in
``oracle/PINNING.md``.
"""
