"""Term lists and symmetry maps of the test and benchmark workloads
(spin operators, ``pauli=0``; PBC bonds)."""

import numpy as np


def chain_maps(L):
    i = np.arange(L)
    return {"T": (i + 1) % L, "P": i[::-1].copy(), "Z": -(i + 1)}


def xxz_static(L, J=1.0, Delta=1.0, h=None, pbc=True):
    nb = L if pbc else L - 1
    hop = [[0.5 * J, i, (i + 1) % L] for i in range(nb)]
    zz = [[J * Delta, i, (i + 1) % L] for i in range(nb)]
    st = [["+-", hop], ["-+", hop], ["zz", zz]]
    if h is not None:
        st.append(["z", [[float(h[i]), i] for i in range(L)]])
    return st


def flat_terms(static):
    """[[opstr, [[J, i, ...], ...]], ...] -> [(opstr, indx, J), ...]."""
    out = []
    for opstr, bonds in static:
        for b in bonds:
            out.append((opstr, tuple(int(x) for x in b[1:]), b[0]))
    return out


def ladder_maps(Lx, Ly):
    i = np.arange(Lx * Ly)
    x, y = i % Lx, i // Lx
    return {"tx": (x + 1) % Lx + y * Lx, "ty": x + ((y + 1) % Ly) * Lx, "px": x[::-1] + y * Lx,
            "py": x + y[::-1] * Lx, "z": -(i + 1)}


def hubbard_static(L, t=1.0, U=4.0, mu=0.0, pbc=True):
    nb = L if pbc else L - 1
    hop_r = [[-t, i, (i + 1) % L] for i in range(nb)]
    hop_l = [[+t, i, (i + 1) % L] for i in range(nb)]
    st = [["+-|", hop_r], ["-+|", hop_l], ["|+-", hop_r], ["|-+", hop_l]]
    if U:
        st.append(["n|n", [[U, i, i] for i in range(L)]])
    if mu:
        st += [["n|", [[-mu, i] for i in range(L)]], ["|n", [[-mu, i] for i in range(L)]]]
    return st
