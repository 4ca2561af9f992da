"""CPU FP64 oracle for the finite-temperature CCSD amplitude iteration.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``kelvin_amd``)
imports, links or runs anything from this directory.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and only as the checker (or the timed CPU baseline), never as the thing
that is measured or shipped.

Modules
-------
``ftcc``        numpy restatement of the reference's hot path (quadrature,
                Fermi factors, UEG integrals, Fermi-weighted blocks, Stanton
                CCSD residual, exp(-tau D) integrators, grand-potential, the
                damped fixed-point solver).
``fock_space``  brute-force Fock-space checks used to pin the residual:
                <Phi_mu| exp(-T) H exp(T) |Phi_0> by explicit matrices, and
                the finite-temperature full-CI grand potential.
"""
