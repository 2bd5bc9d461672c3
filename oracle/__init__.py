"""CPU oracle: a NumPy FP64 restatement of the reference solver's hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package ``hit_solver_amd``
imports this package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker or as
the timed CPU baseline.  The HIP backend never falls back to it.

It restates, function by function, the arithmetic of the reference
(HIT_Solver, pure Python/NumPy) in the same operation order, so that its
results are bit-identical to the reference on the same inputs.  The
reference's algorithm choices that set its CPU cost are kept on purpose
(FFT circulant compact solve, full-grid WENO7 evaluation, Python row-loop
Thomas solve, 80-shift mask dilation), so timing it is timing the
reference's method.

How each part is pinned (tests/test_oracle_pins.py, tests/test_oracle_golden.py):

* golden vectors written from the reference itself (tests/golden/
  make_golden.py -> reference_vectors.npz): sensor masks, rhs, viscous
  terms, hyperviscosity, one SSP-RK3 step of the HIT closure and of the
  Riemann operator, three forcing updates, Mach control, short HIT and
  Riemann runs, the HIT initial condition -- compared bit-exactly;
* closed forms: compact-scheme Fourier symbol, WENO7 invariants (constants,
  linear data -> upwind formula, the IS0 544 quirk), Thomas vs dense solve,
  hyperviscosity Fourier symbol, viscous shear wave, SSP-RK3 stability
  polynomial, initial KE = Mt^2/2, Riemann step counts of the reference's
  own result files.

Modules:
    equations    -- gas algebra (reference OOP/equations.py:125-267)
    periodic     -- periodic sensor / hybrid RHS / HV (OOP/spatial_operator.py)
    nonperiodic  -- Riemann operator (OOP/problems/riemann_config3.py)
    forcing      -- OU shell forcing (OOP/forcing.py)
    hit          -- HIT sources, init, Mach control, step loop (OOP/hit2d.py)
    riemann      -- Riemann configs and loop (OOP/problems/*.py)
    rk           -- SSP-RK3 (OOP/time_operator.py:17-26)
"""
