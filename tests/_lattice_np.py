"""Independent numpy formulations used to pin the oracle (array shifts instead of
per-site index arithmetic)."""
import numpy as np

import oracle as O


def links_to_field(u18: np.ndarray, extent) -> np.ndarray:
    """(nsite*D, 18) reference layout -> complex array indexed [x_{D-1}, ..., x_0, mu, 3, 3]."""
    dim = len(extent)
    z = O.to_np(u18.reshape(-1, dim, 18))
    return z.reshape(tuple(reversed(extent)) + (dim, 3, 3))


def field_to_links(f: np.ndarray) -> np.ndarray:
    dim = f.shape[-3]
    return O.from_np(f.reshape(-1, dim, 3, 3)).reshape(-1, 18)


def shift(a: np.ndarray, mu: int, dim: int, step: int = 1) -> np.ndarray:
    """b[x] = a[x + step * mu_hat] (periodic); axis of direction mu is dim-1-mu."""
    return np.roll(a, -step, axis=dim - 1 - mu)


def dag(m):
    return np.conj(np.swapaxes(m, -1, -2))


def plaquette_field(f: np.ndarray, mu: int, nu: int) -> np.ndarray:
    """P_mu,nu(x) = U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ for every x."""
    dim = f.shape[-3]
    umu = f[..., mu, :, :]
    unu = f[..., nu, :, :]
    return umu @ shift(unu, mu, dim) @ dag(shift(umu, nu, dim)) @ dag(unu)


def hamiltonian_links_np(u18, extent, beta):
    f = links_to_field(u18, extent)
    dim = len(extent)
    s = 0.0
    for mu in range(dim):
        for nu in range(mu + 1, dim):
            p = plaquette_field(f, mu, nu)
            s += np.sum(1.0 - np.trace(p, axis1=-2, axis2=-1).real / 3.0)
    return beta * s


def staple_field(f: np.ndarray, mu: int) -> np.ndarray:
    """Sum over nu != mu of upper and lower staples closing U_mu(x) into plaquettes."""
    dim = f.shape[-3]
    umu = f[..., mu, :, :]
    acc = np.zeros_like(umu)
    for nu in range(dim):
        if nu == mu:
            continue
        unu = f[..., nu, :, :]
        up = shift(unu, mu, dim) @ dag(shift(umu, nu, dim)) @ dag(unu)
        lo_nu = shift(unu, nu, dim, -1)                 # U_nu(x - nu)
        lo = dag(shift(lo_nu, mu, dim)) @ dag(shift(umu, nu, dim, -1)) @ lo_nu
        acc = acc + up + lo
    return acc


GM = [
    np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]], dtype=complex),
    np.array([[0, -1j, 0], [1j, 0, 0], [0, 0, 0]], dtype=complex),
    np.array([[1, 0, 0], [0, -1, 0], [0, 0, 0]], dtype=complex),
    np.array([[0, 0, 1], [0, 0, 0], [1, 0, 0]], dtype=complex),
    np.array([[0, 0, -1j], [0, 0, 0], [1j, 0, 0]], dtype=complex),
    np.array([[0, 0, 0], [0, 0, 1], [0, 1, 0]], dtype=complex),
    np.array([[0, 0, 0], [0, 0, -1j], [0, 1j, 0]], dtype=complex),
    np.diag([1, 1, -2]).astype(complex) / np.sqrt(3),
]
T = [g / 2 for g in GM]


def su3_matrix(v: np.ndarray) -> np.ndarray:
    """(..., 8) real -> (..., 3, 3) sum_a v_a T^a."""
    return np.einsum("...a,aij->...ij", v, np.array(T))


def expi(v: np.ndarray) -> np.ndarray:
    """exp(i v^a T^a) for (..., 8) via eigendecomposition of the hermitian matrix."""
    h = su3_matrix(v)
    w, q = np.linalg.eigh(h)
    return (q * np.exp(1j * w)[..., None, :]) @ dag(q)


def clover_field(f: np.ndarray, i: int, j: int) -> np.ndarray:
    """The four (i, j) plaquettes around x, each a closed loop from x:
    C_ij(x) = U_i(x) U_j(x+i) U_i(x+j)^+ U_j(x)^+
            + U_j(x) U_i(x+j-i)^+ U_j(x-i)^+ U_i(x-i)
            + U_i(x-i)^+ U_j(x-i-j)^+ U_i(x-i-j) U_j(x-j)
            + U_j(x-j)^+ U_i(x-j) U_j(x-j+i) U_i(x)^+"""
    dim = f.shape[-3]
    ui = f[..., i, :, :]
    uj = f[..., j, :, :]
    ui_mi = shift(ui, i, dim, -1)
    uj_mj = shift(uj, j, dim, -1)
    q1 = ui @ shift(uj, i, dim) @ dag(shift(ui, j, dim)) @ dag(uj)
    q2 = uj @ dag(shift(ui_mi, j, dim)) @ dag(shift(uj, i, dim, -1)) @ ui_mi
    q3 = dag(ui_mi) @ dag(shift(uj_mj, i, dim, -1)) @ shift(ui_mi, j, dim, -1) @ uj_mj
    q4 = dag(uj_mj) @ shift(ui, j, dim, -1) @ shift(uj_mj, i, dim) @ dag(ui)
    return q1 + q2 + q3 + q4


def permutation_sign(idx) -> int:
    """Sign of the permutation that sorts idx (0 when a value repeats)."""
    idx = list(idx)
    if len(set(idx)) != len(idx):
        return 0
    inv = sum(1 for a in range(len(idx)) for b in range(a + 1, len(idx)) if idx[a] > idx[b])
    return -1 if inv % 2 else 1
