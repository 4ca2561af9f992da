"""Integral-provider protocol (kelvin/system.py:1-167).

The finite-temperature CCSD solver uses, for a generalised-orbital system:
``verify(T, mu)``, ``has_u()``, ``const_energy()``, ``get_mp1()``,
``g_energies_tot()`` (n orbital energies), ``g_fock_tot()`` (n x n Fock
matrix with its diagonal) and ``g_aint_tot()`` (antisymmetrised
<pq||rs>, n^4).  Arrays may be numpy arrays or torch tensors; the solver
moves them to the GPU once per solve.
"""


class System(object):
    """Base class representing a physical system."""

    def __init__(self):
        pass

    def _missing(self, name):
        raise Exception("Base class function {}()".format(name))

    def verify(self, T, mu):
        """Verify that the system is consistent with the given T, mu."""
        self._missing("verify")

    def has_g(self):
        self._missing("has_g")

    def has_u(self):
        self._missing("has_u")

    def has_r(self):
        self._missing("has_r")

    def const_energy(self):
        """Constant energy shift."""
        self._missing("const_energy")

    def get_mp1(self):
        """First-order grand potential <H - H0>_0."""
        self._missing("get_mp1")

    def g_energies_tot(self):
        """All generalised-orbital energies."""
        self._missing("g_energies_tot")

    def g_fock_tot(self):
        """Generalised-orbital Fock matrix including its diagonal."""
        self._missing("g_fock_tot")

    def g_aint_tot(self):
        """Antisymmetrised generalised-orbital two-electron integrals."""
        self._missing("g_aint_tot")

    def g_mp1_den(self):
        self._missing("g_mp1_den")

    def g_fock_d_den(self):
        self._missing("g_fock_d_den")
