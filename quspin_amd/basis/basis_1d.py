"""``spin_basis_1d`` front-end (src/quspin/basis/basis_1d/spin.py:48-230)
expressed as general-basis maps on the device core.

kblock -> translation by ``a`` sites (period L/a, q = kblock), pblock=+-1 ->
reflection i -> L-1-i (q = 0 / 1), zblock=+-1 -> spin inversion, pzblock ->
reflection with inversion, zAblock / zBblock -> spin inversion of the even /
odd sites.  These are the maps and quantum numbers under which the
reference's own test (test/test_general_spin_opstr.py:58-135,166-250)
requires ``spin_basis_1d`` and ``spin_basis_general`` to agree in states,
normalisations and every ``Op`` matrix element (translation, parity at
k = 0 and k = L/2, spin inversion), so on those sectors this basis is the
1D core's.

Momentum with parity (or parity x inversion) at 0 < k < L/(2a) is the one
sector family the 1D core builds differently: each representative enters
twice, as the real cosine / sine combinations sigma = +-1 of k and -k
(base_1d.py:1001-1087, the ``N < 0`` rows), a construction a group of
commuting maps cannot express.  Those sectors raise NotImplementedError
instead of silently building the general basis' (different) sector.
"""

import warnings

import numpy as np

from .base_general import GeneralBasisWarning
from .spin import spin_basis_general

_EXPECTED = {"kblock", "zblock", "zAblock", "zBblock", "pblock", "pzblock", "a", "check_z_symm", "count_particles",
             "_Np"}


def _q(pm):
    if pm not in (1, -1):
        raise ValueError("blocks take the values +1 or -1")
    return 0 if pm == 1 else 1


class spin_basis_1d(spin_basis_general):
    def __init__(self, L, Nup=None, m=None, S="1/2", pauli=1, Ns_block_est=None, make_basis=True, **blocks):
        wrong = set(blocks) - _EXPECTED
        if wrong:
            raise ValueError("unexpected optional argument(s): " + ", ".join(sorted(wrong)))
        a = blocks.get("a")
        a = 1 if a is None else a
        kblock, pblock, zblock = blocks.get("kblock"), blocks.get("pblock"), blocks.get("zblock")
        pzblock, zAblock, zBblock = blocks.get("pzblock"), blocks.get("zAblock"), blocks.get("zBblock")
        if type(L) is not int or L <= 0:
            raise TypeError("L must be a positive integer")
        if type(a) is not int:
            raise TypeError("a must be integer")
        if L % a != 0:
            raise ValueError("L must be interger multiple of lattice spacing a")
        for name, v in (("pblock", pblock), ("zblock", zblock), ("pzblock", pzblock), ("zAblock", zAblock),
                        ("zBblock", zBblock)):
            if v is not None:
                if type(v) is not int:
                    raise TypeError(f"{name} must be integer")
                if abs(v) != 1:
                    raise ValueError(f"{name} must be +/- 1")
        if pblock is not None and (zAblock is not None or zBblock is not None):
            raise ValueError("zA and zB symmetries incompatible with parity symmetry")
        if blocks.get("count_particles"):
            raise NotImplementedError("count_particles is not supported")
        if blocks.get("_Np") is not None:
            if Nup is not None:
                raise ValueError("do not use _Np and Nup/nb simultaineously.")
            if blocks["_Np"] != -1:
                raise NotImplementedError("_Np (several particle sectors with counts) is not supported")
        nup_given = Nup is not None or m is not None
        if blocks.get("check_z_symm", True) is not False and (zblock is not None or pzblock is not None):
            n_up = Nup if m is None else int((m + 0.5) * L)
            if n_up is not None:
                if L % 2 != 0:
                    raise ValueError("spin inversion/particle-hole symmetry with particle/magnetization conservation "
                                     "must be used with chains with 0 magnetization sector or at half filling")
                if not isinstance(n_up, (int, np.integer)) or int(n_up) % (L + 1) != L // 2:
                    raise ValueError("spin inversion/particle-hole symmetry only reduces the 0 magnetization or half "
                                     "filled particle sector")
        if nup_given and (zAblock is not None or zBblock is not None):
            raise ValueError("zA/cA and zB/cB symmetries incompatible with magnetisation/particle symmetry")
        if kblock is not None and type(kblock) is not int:
            raise TypeError("kblock must be integer")
        nk = L // a
        if kblock is not None:
            kblock = kblock % nk
            if (pblock is not None or pzblock is not None) and (2 * kblock) % nk != 0:
                raise NotImplementedError(
                    "momentum with parity at 0 < k < pi: the 1D core builds the real sigma = +-1 states of k and "
                    "-k there, which commuting symmetry maps cannot express")
        i = np.arange(L)
        maps = {}
        if kblock is not None:
            maps["kblock"] = ((i + a) % L, kblock)
        if pblock is not None:
            maps["pblock"] = (i[::-1].copy(), _q(pblock))
        if zblock is not None:
            maps["zblock"] = (-(i + 1), _q(zblock))
        if pzblock is not None and not (pblock is not None and zblock is not None):
            maps["pzblock"] = (-(i[::-1] + 1), _q(pzblock))
        if zAblock is not None:
            maps["zAblock"] = (np.where(i % 2 == 0, -(i + 1), i), _q(zAblock))
        if zBblock is not None:
            maps["zBblock"] = (np.where(i % 2 == 1, -(i + 1), i), _q(zBblock))
        self._L = L
        # translation and reflection do not commute as maps; at k = 0 and
        # k = L/(2a), the only momenta allowed with parity here, the quantum
        # numbers are invariant, so the general basis' warning does not apply
        with warnings.catch_warnings():
            warnings.filterwarnings("ignore", message="using non-commuting symmetries", category=GeneralBasisWarning)
            spin_basis_general.__init__(self, L, Nup=Nup, m=m, S=S, pauli=pauli, Ns_block_est=Ns_block_est,
                                        make_basis=make_basis, **maps)

    @property
    def L(self):
        return self._L
