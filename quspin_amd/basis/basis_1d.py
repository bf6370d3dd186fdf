"""``spin_basis_1d`` front-end (src/quspin/basis/basis_1d/spin.py) expressed
as general-basis maps on the device core.

kblock -> translation by ``a`` sites (period L/a, q = kblock), pblock=+-1 ->
reflection i -> L-1-i (q = 0 / 1), zblock=+-1 -> spin inversion, pzblock ->
reflection with inversion.  The 1D core's state ordering and phase gauge are
not reproduced: spectra, dynamics and expectation values agree; individual
reduced matrix elements can differ by the gauge of each basis state.
"""

import numpy as np

from .spin import spin_basis_general


def _q(pm):
    if pm not in (1, -1):
        raise ValueError("blocks take the values +1 or -1")
    return 0 if pm == 1 else 1


class spin_basis_1d(spin_basis_general):
    def __init__(self, L, Nup=None, m=None, S="1/2", pauli=1, a=1, kblock=None, pblock=None, zblock=None,
                 pzblock=None, Ns_block_est=None, make_basis=True, **kwargs):
        if kwargs:
            raise NotImplementedError("unsupported spin_basis_1d arguments: {}".format(sorted(kwargs)))
        i = np.arange(L)
        blocks = {}
        if kblock is not None:
            if L % a:
                raise ValueError("L must be a multiple of a")
            blocks["kblock"] = ((i + a) % L, int(kblock) % (L // a))
        if pblock is not None:
            blocks["pblock"] = (i[::-1].copy(), _q(pblock))
        if zblock is not None:
            blocks["zblock"] = (-(i + 1), _q(zblock))
        if pzblock is not None:
            blocks["pzblock"] = (-(i[::-1] + 1), _q(pzblock))
        self._L = L
        spin_basis_general.__init__(self, L, Nup=Nup, m=m, S=S, pauli=pauli, Ns_block_est=Ns_block_est,
                                    make_basis=make_basis, **blocks)

    @property
    def L(self):
        return self._L
