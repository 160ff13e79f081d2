"""CPU oracle for the force-and-integrate path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import, call, link or run anything in this directory, and only as the checker or as
the timed CPU baseline. The product path (``mdpy_amd``) never routes through it.

Contents
--------
``mdref``   float64 numpy restatement of mdpy v0.2.1's CPU kernels and integrators
            (wrap / minimum image, cell list, CHARMM LJ, plain Coulomb, bond, angle + UB,
            dihedral, improper, Ensemble.update, Verlet, Langevin), each function citing the
            reference file:line it follows; plus exact (non-defective) integrators.
``ewald``   float64 smooth-PME (Essmann et al., J. Chem. Phys. 103, 8577 (1995)),
            brute-force Ewald sums, erfc direct space, self / exclusion / background terms.
            The reference has no PME; these define the new constraint's physics.
``refpath.c`` C restatement of the reference CPU step (27-cell double-visit LJ, all-pairs
            Coulomb, serial bonded loops, cell-list rebuild per position update, three force
            evaluations per Langevin step) used only as the timed CPU baseline.

Pinning (what checks each function, beyond the function's own formula):
    wrap/unwrap            mdpy/test/test_pbc.py:26-81 vectors
    cell list              mdpy/test/test_cell_list.py:41-58 on 6PO6.pdb; brute-force binning
    neighbour sets         scipy cKDTree periodic pair search (independent algorithm)
    LJ                     mdpy/test/test_charmm_nonbonded_constraint.py:93-127 KAT; central FD
    Coulomb                mdpy/test/test_electrostatic_constraint.py:197-221 KAT; central FD
    bond / angle / improper  KATs of test_charmm_{bond,angle,improper}_constraint.py; FD
    dihedral               test_charmm_dihedral_constraint.py KAT (energy, net force, torque);
                           FD pins the corrected force; the reference force is pinned only by
                           its zero net force/torque (its magnitude is never tested upstream)
    SPME                   brute-force Ewald (converged), NaCl Madelung constant 1.747565,
                           beta-independence of the total, FD of the reciprocal energy
    integrators            energy conservation (VV), exact harmonic-oscillator trajectory,
                           reference update algebra on a 1-atom free particle (closed form)
"""
