"""CPU oracle for the event-batched SNR / Fisher path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import this package, and only as
the checker (or the timed CPU baseline); the product path in ``gwfast_amd``
never calls into it and fails loudly when its HIP library is missing.

What it is: an FP64 NumPy restatement of gwfast v1.1.2's algorithm for this
path, written from the reference source read as text (every function cites
the reference file:line it follows).  JAX's reverse-mode AD, which the
reference uses for the Mc, eta, chi1z, chi2z (and LambdaTilde, deltaLambda)
rows (gwfast/signal.py:1143-1189), is replaced by forward-mode dual numbers
(``oracle.dual``): both produce the exact derivative of the same FP64
expression graph, differing only by rounding.

How it is pinned (tests/test_oracle_*.py):
  * notebook known answers: TaylorF2 Schwarzschild/Kerr cut 1590.1/3433.5 Hz
    (notebooks/new_features_tutorial.ipynb:115), ET optimal location
    [0.8636426, 0.1643505] (notebooks/gwfast_tutorial.ipynb:662-671);
  * closed forms / textbook routines: non-spinning 3.5PN TaylorF2 phase and
    tau(f) in canonical-fraction form, detector-tensor pattern functions,
    unit-vector light-travel delay, F[dL,dL] = SNR^2/dL^2, F[Phic,Phic] =
    SNR^2, trapezoid sums of power laws, mpmath matrix inverse / eigenvalues;
  * invariants: IMRPhenomD C1 continuity of the phase and C0 continuity of
    the amplitude at the joins, triangle arms summing to zero, F+^2+Fx^2
    independent of psi;
  * derivatives: dual numbers vs complex step (TaylorF2 is holomorphic) and
    vs central finite differences; the reference's closed-form derivative
    rows (signal.py:1432-1577) vs dual numbers;
  * fit-coefficient tables of IMRPhenomD compared number by number with the
    reference source text (CPU test, skipped where the reference is absent).

Parity unpinned (no reference output exists anywhere in this pipeline, JAX
is not installed): XLA's transcendental rounding, jnp.geomspace interior
rounding, and the reduction order of jnp.trapezoid.  The stated tolerances
(SNR <= 1e-10 relative, Fisher <= 1e-6 normalised) absorb these.
"""
