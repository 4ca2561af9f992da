"""Block containers with the field names of cqcpy's ov_blocks (the types the
reference passes around as F and I, kelvin/cc_utils.py:588,599-601).

``one_e_blocks`` / ``two_e_blocks`` hold user arrays (numpy or torch).
``EngineBlocks`` stands for blocks that live inside an FTEngine: reading a
field exports a device copy; the hot path uses the engine's own buffers.
"""
from .engine import ONE_E, TWO_E


class one_e_blocks(object):
    def __init__(self, oo=None, ov=None, vo=None, vv=None):
        self.oo, self.ov, self.vo, self.vv = oo, ov, vo, vv


class two_e_blocks(object):
    def __init__(self, vvvv=None, vvvo=None, vovv=None, vvoo=None, vovo=None,
                 oovv=None, vooo=None, ooov=None, oooo=None):
        self.vvvv, self.vvvo, self.vovv = vvvv, vvvo, vovv
        self.vvoo, self.vovo, self.oovv = vvoo, vovo, oovv
        self.vooo, self.ooov, self.oooo = vooo, ooov, oooo


class EngineBlocks(object):
    """Handle on the Fermi-weighted blocks stored in ``engine``."""

    def __init__(self, engine, names):
        object.__setattr__(self, "engine", engine)
        object.__setattr__(self, "_names", tuple(names))
        object.__setattr__(self, "_cache", {})

    def __getattr__(self, name):
        if name not in self._names:
            raise AttributeError(name)
        if name not in self._cache:
            t = self.engine.get_block(name)
            t._ftcc_source = (self.engine, name)
            self._cache[name] = t
        return self._cache[name]


def engine_fock(engine):
    return EngineBlocks(engine, ONE_E)


def engine_eri(engine):
    return EngineBlocks(engine, TWO_E)


class EngineDenominators(object):
    """The engine forms D1/D2 from its orbital energies on the fly."""

    def __init__(self, engine, which):
        self.engine, self.which = engine, which
