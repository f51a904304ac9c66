"""Closed-form gauge fields whose field strength and clover term are known
exactly, used to pin the clover-leaf field strength (the oracle's and the
device's) independently of either implementation.

Constant abelian flux through one (mu, nu) plane: with T = diag(1, -1, 0) and
B = 2 pi n / (L_mu L_nu),
    U_nu(x) = exp(i B x_mu T),
    U_mu(x) = exp(-i B L_mu x_nu T) on the slice x_mu = L_mu - 1, 1 elsewhere,
every plaquette of the plane, at every site and in either orientation of the
clover leaves of tests/mesfield.cc:24-64, equals exp(i B T) (the jump of the
U_nu phase across the x_mu wrap is undone by the U_mu phase on that slice, and
B L_mu L_nu is a multiple of 2 pi).  All links are diagonal and commute, so
the four leaves agree and
    F_{mu nu} = (Q - Q^+) / 8 = i sin(B T)   (= i sin(B) T for T = diag(1, -1, 0)),
while the other planes carry no flux.  A leaf traversed backwards gives
half of it and a wrong normalisation a wrong multiple, so either fails a
comparison with this value.  Fluxes through two disjoint planes (e.g. 01 and
23) superpose without touching the remaining four planes, and a constant
SU(3) rotation W (U -> W U W^+) turns T into a full Hermitian matrix, so
every colour entry of F and every E / B entry of the clover packing is
exercised.
"""
import numpy as np

PLANES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def site_coords(L):
    """Coordinates of the canonical (QDP++ checkerboarded) site order
    i = cb*Vh + ((t*Nz + z)*Ny + y)*Nxh + x/2."""
    V = L[0] * L[1] * L[2] * L[3]
    vh, nxh = V // 2, L[0] // 2
    i = np.arange(V)
    cb, j = i // vh, i % vh
    xh = j % nxh
    r = j // nxh
    y = r % L[1]
    r //= L[1]
    z = r % L[2]
    t = r // L[2]
    x = 2 * xh + ((y + z + t + cb) & 1)
    return np.stack([x, y, z, t], axis=1)


def random_su3(seed):
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    Q, R = np.linalg.qr(G)
    Q = Q @ np.diag(np.diag(R) / np.abs(np.diag(R)))
    return Q / np.linalg.det(Q) ** (1.0 / 3.0)


def abelian_flux(L, fluxes, rot_seed=None):
    """Gauge field with constant flux n through each plane in `fluxes`
    ({(mu, nu): (n, T_diag)}, disjoint planes), optionally rotated by a
    constant SU(3) matrix.  Returns (u, F): u canonical (V*4*18 doubles,
    U_mu(x) row-major), F[plane] the 3x3 field strength (same on all sites)."""
    x = site_coords(L)
    V = x.shape[0]
    phase = np.zeros((V, 4, 3))
    Fk = {p: np.zeros((3, 3), dtype=np.complex128) for p in PLANES}
    W = random_su3(rot_seed) if rot_seed is not None else np.eye(3)
    for (mu, nu), (n, tdiag) in fluxes.items():
        T = np.asarray(tdiag, dtype=np.float64)
        B = 2.0 * np.pi * n / (L[mu] * L[nu])
        phase[:, nu, :] += B * x[:, mu, None] * T[None, :]
        edge = x[:, mu] == L[mu] - 1
        phase[edge, mu, :] += -B * L[mu] * x[edge, nu, None] * T[None, :]
        Fk[(mu, nu)] = W @ np.diag(1j * np.sin(B * T)) @ W.conj().T
    U = np.zeros((V, 4, 3, 3), dtype=np.complex128)
    U[:, :, range(3), range(3)] = np.exp(1j * phase)
    U = np.einsum("ab,nmbc,dc->nmad", W, U, W.conj())
    return np.ascontiguousarray(U).view(np.float64).ravel().copy(), Fk


def gammas_degrand_rossi():
    """QDP++ DeGrand-Rossi gamma matrices (gamma_5 = diag(1, 1, -1, -1))."""
    i = 1j
    g0 = np.array([[0, 0, 0, i], [0, 0, i, 0], [0, -i, 0, 0], [-i, 0, 0, 0]])
    g1 = np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]])
    g2 = np.array([[0, 0, i, 0], [0, 0, 0, -i], [-i, 0, 0, 0], [0, i, 0, 0]])
    g3 = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]])
    return [np.asarray(g, dtype=np.complex128) for g in (g0, g1, g2, g3)]


def clover_dense_closed_form(Fk, csw_r, csw_t, diag_mass):
    """12x12 clover matrix (index 3*spin + colour) of the Sheikholeslami-
    Wohlert term with leaf-average F (anti-Hermitian) and
    sigma_{mu nu} = [g_mu, g_nu] / 2:
        A = diag_mass - sum_{mu<nu} (c_{mu nu} / 2) sigma_{mu nu} (x) F_{mu nu},
    c = csw_t on the planes containing t, csw_r on the others."""
    g = gammas_degrand_rossi()
    A = diag_mass * np.eye(12, dtype=np.complex128)
    for (m, n) in PLANES:
        sig = 0.5 * (g[m] @ g[n] - g[n] @ g[m])
        c = csw_t if n == 3 else csw_r
        A -= 0.5 * c * np.kron(sig, Fk[(m, n)])
    return A


def unpack_triangular(site72):
    """Chroma PrimitiveClovTriang site record (diag[2][6], offd[2][15][2]) ->
    two dense Hermitian 6x6 blocks."""
    out = []
    for b in range(2):
        d = site72[6 * b:6 * b + 6]
        o = site72[12 + 30 * b:12 + 30 * b + 30].view(np.complex128)
        M = np.diag(d).astype(np.complex128)
        for i in range(6):
            for j in range(i):
                M[i, j] = o[i * (i - 1) // 2 + j]
                M[j, i] = np.conj(M[i, j])
        out.append(M)
    return out


def pack_triangular(blocks):
    s = np.zeros(72)
    for b, M in enumerate(blocks):
        s[6 * b:6 * b + 6] = M.diagonal().real
        o = np.zeros(15, dtype=np.complex128)
        for i in range(6):
            for j in range(i):
                o[i * (i - 1) // 2 + j] = M[i, j]
        s[12 + 30 * b:12 + 30 * b + 30] = o.view(np.float64)
    return s
