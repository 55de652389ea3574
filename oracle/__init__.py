"""CPU oracle for the closed-shell PPP CCSD (``r-ccsd``) path of helios.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``helios_amd``, the HIP
library, the Fortran host) may import, link or call anything in this package.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker.

Contents (every function cites the reference file:line it restates; paths are
relative to the helios source tree):

* ``ppp``      -- PPP model build: Ohno-Klopman / Mataga-Nishimoto / Hubbard
                  gamma, core Hamiltonian, Hueckel matrix
                  (src/input/readMoleculeInformation.f90).
* ``scf``      -- closed-shell RHF by the reference's damped-density Roothaan
                  iteration (src/methods/mod_scf.f90), MO Fock and reference
                  energy (src/methods/coupledCluster/mod_cc.f90:2311-2504).
* ``ccsd``     -- the spatial-orbital closed-shell CCSD of the reference,
                  restated term by term: integrals R = 2(pq|rs)-(ps|rq),
                  singles/doubles projections (ccsd/proj-ccsd-spat-hf-1/2.f90),
                  guess, step, DIIS, energy, iterator statuses.
* ``spin_ccsd``-- an independent spin-orbital CCSD (Stanton, Gauss, Watts,
                  Bartlett, J. Chem. Phys. 94, 4334 (1991)) used only to pin
                  the spatial oracle.
* ``fci``      -- determinant-basis full CI for tiny molecules (pins the
                  two-electron case where CCSD is exact).
* ``dressed``  -- the T1-dressed, ZDO-factorised residual that the GPU kernels
                  implement, written as numpy GEMMs; pinned against ``ccsd``.
* ``mlp``      -- the learned T2 guess (features, BF16-rounded MLP forward,
                  analytic backward, Adam) restated in float64 numpy; no
                  reference counterpart, pinned by finite differences.

The reference itself (Fortran, Intel-compiler extensions, ~20 modules with
generated includes) is not built here: it has no stand-alone source set for
this path that AMD flang compiles without stand-in modules, so there is no
``oracle/_ref``.  ``bench.py`` times ``refloops`` (a C restatement of the
reference's loop nests, ``oracle/refloops.c``) on the host cores as the CPU
baseline ("port").

Pinning status of each function is recorded in the module docstrings.
"""
