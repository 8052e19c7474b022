"""CPU restatement of the BayesianGroupTesting hot path -- TEST INFRASTRUCTURE ONLY.

This package is the parity checker for the HIP provider and the timed CPU
baseline. Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it. The provider package never imports it.

Each restated function is pinned by the reference's own known-answer tests
(``tests/golden/reference_kats.json``, transcribed with citations) and by checks the
mathematics fixes (closed forms, brute force, planted mistakes that must fail):
``tests/test_oracle_kats.py`` and ``tests/test_oracle_properties.py``.
"""
