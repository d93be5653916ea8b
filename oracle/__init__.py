"""CPU oracle for the BASS sparse-state hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy (and, under ``oracle/port/``, in single-thread
C++), the algorithm of the reference's sparse-state core:

* ``kernels.py``  -- the Numba kernels of ``src/core/sparse_state.py``,
  ``src/core/truncation.py`` and ``src/simulation/bass_simulator.py``;
* ``drivers.py``  -- ``SparseState``, ``BASS`` and ``FixedBasisSimulator``
  (``src/core/sparse_state.py:193-415``, ``src/simulation/bass_simulator.py:478-929``,
  ``src/simulation/simulator.py:61-253``);
* ``exact.py``    -- the dense ``ExactSimulator`` and the fidelity / participation
  ratio metrics (``src/simulation/exact_simulator.py``, ``src/benchmarking/fidelity.py``);
* ``circuits.py`` -- gate matrices and circuit generators (``src/core/gates.py``,
  ``src/utils/random_circuits.py``, ``src/experimentation/runner.py``).

Every function cites the reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import anything from here, and only as the checker.  The product
(the HIP library and its Python host layer) never imports, links or executes
this package.

Parity pins (what checks each function; the tests named live in ``tests/``):

* gate matrices / TFIM + RFIM generators / dense simulator: closed forms and the
  reference-held ``rotation_diag.csv`` PRZ value (``test_oracle_fixtures.py``);
* 1q / 2q emit, merge, radix sort, diagonal, hash 2q, rotation, RDMs, eigh_2x2:
  dense brute force at n <= 10, ``numpy.linalg.eigh``, invariants
  (``test_oracle_kernels.py``);
* BASS / FixedBasis drivers end to end: the reference-held per-trial fidelity and
  participation ratio in ``tests/results/fidelity_pr_raw.npz`` (RFIM rows) and
  ``nscaling_*.npz`` (``test_oracle_fixtures.py``);
* RFIM and QAOA generators plus the dense simulator: the exact-state PR and
  half-chain entropy of every trial of ``phase_diagram_extended.npz``; the RFIM
  driver rows there as 20-trial means; its Haar families (brickwork 1D / 2D,
  random pairing; tie-free) per trial for both drivers, optimiser included;
  the 2-qubit brick-wall pass end to end against ``gamma_diagnostics.npz``'s
  tie-free rows (``test_oracle_fixtures.py``);
* random-k truncation (LCG partial Fisher-Yates): closed-form LCG sequence.

Parity unpinned (no reference-held result or closed form decides them):

* the reference's order among exactly tied |a|^2 (``np.argpartition`` inside
  and outside Numba): ``argpartition_selector`` reproduces numpy's choice on
  this host only, and numpy's own choice differs between its SIMD dispatch
  paths; single trials that truncate through ties are compared as means;
* the FixedBasis random-k draw (``np.random.choice`` on the legacy generator;
  its only stored results, ``schmidt_benchmark.npz``, were not replayed).
"""
