"""CPU FP64 oracle for the pseudo-spectral Navier-Stokes hot path.

TEST INFRASTRUCTURE ONLY.  This package is a line-by-line numpy restatement of
the reference solver (agdestein/SymmetryCode.jl, Julia) and exists to check the
HIP implementation.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
the timed CPU baseline -- never as a code path of the product.

Layout convention: a Julia column-major array of size ``(n/2+1, n, n)`` (kx
fastest) is held here as a C-order numpy array of shape ``(n, n, n/2+1)``
indexed ``[z, y, x]``; the two have identical memory order.  2D fields are
``[y, x]``.  Vector fields stack their components on a leading axis
(``[x, y, z]``), symmetric tensors as ``[xx, yy, zz, xy, yz, zx]`` (3D) or
``[xx, yy, xy]`` (2D), exactly the reference's NamedTuple order.

Pinning (what checks each function; see ``tests/test_oracle.py``):

* FFT scaling / ``to_spec`` / ``to_phys``: brute-force DFT at n <= 8, the
  reference KAT ``spec[1] == 1.5`` for a constant field, Parseval.
* ``twothirds``, ``cutoff_index``: brute-force loops mirroring the reference
  test (test/test_solver_kernels.jl:6-17).
* ``project``: divergence below 1e-10 (test/test_solver_kernels.jl:19-28) and
  idempotence.
* ``convectiondiffusion``: energy-budget identity, octahedral and translation
  equivariance, closed-form decay of 2D Taylor-Green and a 3D shear wave.
* ``wray3`` / ``rk4``: exact stability polynomials on Taylor-Green.
* closures: the reference KATs (test/test_closures.jl) plus 3D analogues.
* ``gaussianfilter``: closed form with the |k|^2 < 9 carve-out.
* ``band forcing``: injection equals the requested rate exactly.
* ``randomfield``: shell energies equal the prescribed profile.
* ``pmlp`` (pointwise MLP closure): forward and gradients against torch
  autograd, fullchain's degree-two homogeneity, the Re_Delta shift
  (``tests/test_pmlp_oracle.py``).

Functions with no independent check are marked "parity unpinned" in their
docstrings.
"""
