"""Spin Hamiltonians of the benchmark workloads as PauliExpressions (Pauli units).

transverse-field Ising: H = -J sum_<ij> Z_i Z_j - h sum_i X_i
Heisenberg:             H =  J sum_<ij> (X_i X_j + Y_i Y_j + Z_i Z_j)
No sign rule (Marshall rotation) is applied.
"""
from .pauli import sigma_x, sigma_y, sigma_z, PauliExpression


def chain_bonds(N, periodic=True):
    bonds = [(i, i + 1) for i in range(N - 1)]
    if periodic and N > 2:
        bonds.append((N - 1, 0))
    return bonds


def square_bonds(L, periodic=True):
    Lx, Ly = (L, L) if isinstance(L, int) else L
    idx = lambda r, c: r * Ly + c
    bonds = []
    for r in range(Lx):
        for c in range(Ly):
            if c + 1 < Ly or (periodic and Ly > 2):
                bonds.append((idx(r, c), idx(r, (c + 1) % Ly)))
            if r + 1 < Lx or (periodic and Lx > 2):
                bonds.append((idx(r, c), idx((r + 1) % Lx, c)))
    return bonds


def transverse_ising(bonds, N, J=1.0, h=1.0):
    H = PauliExpression()
    for i, j in bonds:
        H = H + (-J) * sigma_z(i) * sigma_z(j)
    for i in range(N):
        H = H + (-h) * sigma_x(i)
    return H


def heisenberg(bonds, J=1.0):
    H = PauliExpression()
    for i, j in bonds:
        H = H + J * (sigma_x(i) * sigma_x(j) + sigma_y(i) * sigma_y(j) + sigma_z(i) * sigma_z(j))
    return H


def transverse_ising_chain(N, J=1.0, h=1.0, periodic=True):
    return transverse_ising(chain_bonds(N, periodic), N, J, h)


def heisenberg_chain(N, J=1.0, periodic=True):
    return heisenberg(chain_bonds(N, periodic), J)


def heisenberg_square(L, J=1.0, periodic=True):
    return heisenberg(square_bonds(L, periodic), J)


def transverse_ising_square(L, J=1.0, h=1.0, periodic=True):
    n = L * L if isinstance(L, int) else L[0] * L[1]
    return transverse_ising(square_bonds(L, periodic), n, J, h)
