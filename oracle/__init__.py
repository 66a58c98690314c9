"""CPU FP64 oracle for the implicit P1-FEM Richards time step.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy/SciPy restatement of the
reference's algorithm (NourEddine-Toutlini/JAX-WSPM), written from its source
text.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
of ``bench.py`` may import it, and only as the checker or the timed CPU
baseline.  The product package ``jax_wspm_amd`` never imports it and never
falls back to it.

Modules and what pins each one (the pin is the test that fails on a plausible
mistake, not a restatement of the formula):

* ``mesh``      structured generators + CSV/zip loader.  Pinned: generated
                arrays equal the reference's ``data/mesh`` fixtures exactly
                (checksums in ``tests/golden``).
* ``soil``      van Genuchten-Mualem and Gardner models.  Pinned: C = dtheta/dh
                by complex-step; saturated limits; Gardner end-to-end by the
                published Test1-3 error metrics and the 3D NPZ.
* ``assembly``  P1 triangle/tet element matrices, COO->CSR, Dirichlet rows.
                Pinned: closed forms (stiffness row sums 0, lumped mass A/3 and
                1.5 V), brute-force dense assembly, end-to-end fixtures.
* ``krylov``    CG / BiCGSTAB / GMRES(restart) with Jacobi in the form of
                ``jax.scipy.sparse.linalg`` (the reference calls it at
                src/solvers/jax_linear_solver.py:282,305,329).  Pinned: dense
                ``numpy.linalg.solve`` on small systems; 3D NPZ runs.
* ``picard``    2D / 3D Picard + adaptive time loops.  Pinned: the reference's
                published Test1-3 metrics (resultsRE2D/*.csv) and the 3D
                trajectories (results3D/*.npz).
* ``tracy``     Tracy analytic solution and the error metrics.  Pinned: the
                published metrics reproduce from the oracle trajectory.
* ``transport`` solute / nitrogen operators.  Parity unpinned against the
                reference (no numeric output of that path is committed there);
                pinned only by invariants and brute force.
"""
