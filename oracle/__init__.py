"""CPU oracle for the trex_amd matching engine -- TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, the reference algorithms the GPU kernels
must match bit-for-bit (integer paths) or within a stated tolerance
(probabilistic path). It is imported only by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``, and
only as the checker: nothing in ``trex_amd`` imports or calls it, and no
product path falls back to it.

Modules and what pins each of them (the tests named here hold the checks):

* ``oracle.dfa`` -- ``DFA.simulate`` (src/regex_to_dfa.py:224-233), the
  byte-table walk of ``cpu_prefix_compose`` / ``cpu_monoid_batch``
  (cuda/prefix_compose.cu:488-516, cuda/monoid_batch.cu:679-699),
  ``simulate_monoid`` (src/monoid.py:174-192) and ``simulate_kgram_monoid``
  (src/kgram.py:195-232). Pinned by the reference's golden vectors
  (tests/golden/reference_vectors.json), by Python ``re.fullmatch`` and by
  closed forms (parity of a's, value mod 3, suffix tests):
  tests/test_oracle.py; the k-gram walk by the same closed forms for every
  k in 2..4 and tail length 0..k-1, and a k-gram table composed in the wrong
  order is shown to fail: tests/test_construction.py.
* ``oracle.nfa`` -- ``simulate_nfa`` (src/nfa_matrices.py:133-158) and a
  set-of-states walk over the packed successor masks. Pinned by the reference's
  ">= 3 a's" golden (cuda/nfa_tc_evolution.cu:1060-1120), by ``re.fullmatch``
  and by agreement of the two independent formulations: tests/test_oracle.py.
* ``oracle.scan`` -- the decoupled look-back composition of
  ``monoid_r3_kernel`` (cuda/monoid_scan.cu:128-222) over DFA functions.
  Pinned by equality with the sequential walk (associativity) and the even-a
  closed form: tests/test_oracle.py.
* ``oracle.prob`` -- FP64 forward pass of the probabilistic NFA. The reference
  has no likelihood (its V5 kernel only thresholds a dot product,
  cuda/nfa_tc_evolution.cu:552-559), so this function is "parity unpinned"
  against the reference; it is pinned by closed forms (identity matrices,
  a single symbol's stationary limit, scipy-free brute-force products):
  tests/test_oracle.py.
* ``oracle.cpu`` -- ctypes wrapper of ``liboracle_cpu.so`` (``cpu_engine.c``),
  a multithreaded C restatement of the DFA byte walk and the NFA set walk,
  used for full-size checks and as the timed CPU baseline. Pinned against the
  Python restatements above: tests/test_oracle.py.
"""
