"""Linear-regime helpers shared by the growth-rate tests (CPU oracle and GPU).

Around the unperturbed state (ee = G0, bee = G0b, mm = bmm = 0) the coherences
x = (emr, emi, bemr, bemi) obey x' = M x with M built from
W[v,k] = w_k (1 - v.v_k) (the INP projection, nuosc_class.cpp:187) and
C0 = W (G0b - G0):
    emr'  = -mu (G0  B + C0 emi),   emi'  = mu (-G0  A + C0 emr)
    bemr' = -mu (G0b B + C0 bemi),  bemi' = mu (-G0b A + C0 bemr)
with A = W (bemr - emr), B = W (bemi + emi) (linearising
nuosc_class.cpp:190-200).  A z-Fourier mode e^{i m 2 pi z / Lz} adds the
semi-discrete symbols of the 4th-order advection and the KO term on the
diagonal.  The largest real part of the spectrum is the growth rate of
|rho_ex| for that mode.
"""
import numpy as np


def linear_operator(vg, G0, G0b, mu):
    vx, vy, vz, vw = vg
    n = len(vw)
    W = vw[None, :] * (1 - (vx[:, None] * vx[None, :] + vy[:, None] * vy[None, :] + vz[:, None] * vz[None, :]))
    C0 = W @ (G0b - G0)
    Z = np.zeros((n, n))
    g, gb = G0[:, None] * W, G0b[:, None] * W
    return np.block([
        [Z, -mu * g - mu * np.diag(C0), Z, -mu * g],
        [mu * g + mu * np.diag(C0), Z, -mu * g, Z],
        [Z, -mu * gb, Z, -mu * gb - mu * np.diag(C0)],
        [mu * gb, Z, -mu * gb + mu * np.diag(C0), Z]])


def mode_growth_rate(vg, G0, G0b, mu, m, nz, dz, ko):
    """Largest real part of the spectrum for z-mode m on an nz-cell box."""
    vz = vg[2]
    th = 2 * np.pi * m / nz
    sym = -1j * vz * (8 * np.sin(th) - np.sin(2 * th)) / (6 * dz) - (ko / dz) * np.sin(th / 2) ** 4
    M = linear_operator(vg, G0, G0b, mu).astype(complex) + np.diag(np.tile(sym, 4))
    return np.linalg.eigvals(M).real.max()


def mode_state(G0, G0b, Z, m, eps, seed=1):
    """Unperturbed state plus a z-mode m coherence with a random angular
    profile: shape (8, 1, nz, nv)."""
    nv, nz = len(G0), len(Z)
    rng = np.random.default_rng(seed)
    Lz = nz * (Z[1] - Z[0])
    kz = 2 * np.pi * m / Lz
    c, s = np.cos(kz * Z)[:, None], np.sin(kz * Z)[:, None]
    pr, pi, qr, qi = (rng.uniform(-1, 1, nv)[None, :] for _ in range(4))
    st = np.zeros((8, 1, nz, nv))
    st[0, 0] = G0[None, :]
    st[4, 0] = G0b[None, :]
    st[2, 0] = eps * G0 * (pr * c - pi * s)
    st[3, 0] = eps * G0 * (pr * s + pi * c)
    st[6, 0] = eps * G0b * (qr * c - qi * s)
    st[7, 0] = eps * G0b * (qr * s + qi * c)
    return st


def fitted_rate(times, rex2):
    """Slope of log|rho_ex| = 0.5 log(sum w |rho_ex|^2) against time."""
    return np.polyfit(np.asarray(times), 0.5 * np.log(np.asarray(rex2)), 1)[0]


def mode_eigenstate(vg, G0, G0b, mu, m, Z, dz, ko, eps):
    """Unperturbed state plus the fastest-growing eigenvector of z-mode m
    (real part of X e^{ikz}), scaled to max |coherence| = eps * max G0.
    Returns (state (8, 1, nz, nv), growth rate)."""
    nv, nz = len(G0), len(Z)
    vz = vg[2]
    th = 2 * np.pi * m / nz
    sym = -1j * vz * (8 * np.sin(th) - np.sin(2 * th)) / (6 * dz) - (ko / dz) * np.sin(th / 2) ** 4
    M = linear_operator(vg, G0, G0b, mu).astype(complex) + np.diag(np.tile(sym, 4))
    ev, vec = np.linalg.eig(M)
    top = np.argmax(ev.real)
    X = vec[:, top]
    X = X / np.max(np.abs(X)) * eps * np.max(G0)
    kz = 2 * np.pi * m / (nz * dz)
    ph = np.exp(1j * kz * Z)[:, None]
    st = np.zeros((8, 1, nz, nv))
    st[0, 0] = G0[None, :]
    st[4, 0] = G0b[None, :]
    for q, f in enumerate((2, 3, 6, 7)):
        st[f, 0] = (X[q * nv:(q + 1) * nv][None, :] * ph).real
    return st, ev.real[top]
