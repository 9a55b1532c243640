"""CPU oracle of the thermal-synchrotron surrogate. TEST INFRASTRUCTURE ONLY (imported by tests/ as the
checker of the GPU kernels in spacetime-optics_amd/csrc/mlp.hip; never part of the product path).

- exact_lnf: the exact normalised coefficients (ln f_j, ln f_a) restated from
  blackholes/src/emission.cuh:159-170 (Leung et al. 2011 emissivity with the K2(1/Theta) table of
  blackholes/src/scene_build.cpp:159-170 read through the float table_lerp of
  blackholes/src/device_util.cuh:74-80, emulated in float32), f_a by Kirchhoff's law in the
  Rayleigh-Jeans limit.
- leung_lnf_scipy: the same formula with K2 from scipy.special.kv (no table), the independent pin.
- forward/backward: FP64 3 -> 64 -> 64 -> 64 -> 2 SiLU network with the device parameter layout.
"""
import numpy as np

NIN, NH, NOUT = 3, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + NIN * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * NH
OFF_W4 = OFF_B3 + NH
OFF_B4 = OFF_W4 + NH * NOUT
NPARAM = OFF_B4 + NOUT
assert NPARAM == 8706

LN_FJ_CONST = np.log(np.sqrt(2.0) * np.pi / 3.0 * 2.0 / 9.0)
FLOOR_J, FLOOR_A = -40.0, -70.0
TH_LO, TH_HI = 0.5, 100.0
LX_LO, LX_HI = -17.0, 14.0       # ln X box of the surrogate coordinates
LS_LO = np.log(1e-3)             # ln sin theta_B >= ln 1e-3


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NIN, NH), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(NH, NH), p[OFF_B3:OFF_W4], p[OFF_W4:OFF_B4].reshape(NH, NOUT), p[OFF_B4:])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, xn):
    """Normalised outputs y[n][2] and the cache for backward."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    x = np.asarray(xn, dtype=np.float64)
    z1 = x @ W1 + b1
    h1 = silu(z1)
    z2 = h1 @ W2 + b2
    h2 = silu(z2)
    z3 = h2 @ W3 + b3
    h3 = silu(z3)
    y = h3 @ W4 + b4
    return y, (x, z1, h1, z2, h2, z3, h3)


def loss_and_grad(p, norm, xn, t_raw, w):
    """Weighted MSE on normalised outputs (loss = sum_i w_i |y_i - t_i|^2 / n) and its gradient."""
    mu = np.array([norm[6], norm[7]], dtype=np.float64)
    sd = np.array([norm[8], norm[9]], dtype=np.float64)
    t = (np.asarray(t_raw, dtype=np.float64) - mu) / sd
    w = np.asarray(w, dtype=np.float64)
    n = len(w)
    y, (x, z1, h1, z2, h2, z3, h3) = forward(p, xn)
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    d = y - t
    loss = np.sum(w[:, None] * d * d) / n
    dy = 2.0 * w[:, None] * d / n
    g = np.zeros(NPARAM)
    g[OFF_W4:OFF_B4] = (h3.T @ dy).ravel()
    g[OFF_B4:] = dy.sum(0)
    dz3 = (dy @ W4.T) * dsilu(z3)
    g[OFF_W3:OFF_B3] = (h2.T @ dz3).ravel()
    g[OFF_B3:OFF_W4] = dz3.sum(0)
    dz2 = (dz3 @ W3.T) * dsilu(z2)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * dsilu(z1)
    g[OFF_W1:OFF_B1] = (x.T @ dz1).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return loss, g


def surrogate_coords(u0, theta, theta_b):
    """(ln X clamped to LX_HI, ln Theta, ln sin theta_B with sin >= 1e-3), X = 10^u0 / ((2/9) Theta^2 sin)."""
    lth = np.log(np.asarray(theta, dtype=np.float64))
    ls = np.log(np.maximum(np.sin(np.asarray(theta_b, dtype=np.float64)), 1e-3))
    lx = np.asarray(u0, dtype=np.float64) * np.log(10.0) - np.log(2.0 / 9.0) - 2.0 * lth - ls
    return np.minimum(lx, LX_HI), lth, ls


def default_input_norm():
    """Input centres[3] and half-widths[3] of the surrogate coordinates' box."""
    lo = np.array([LX_LO, np.log(TH_LO), LS_LO])
    hi = np.array([LX_HI, np.log(TH_HI), 0.0])
    return 0.5 * (lo + hi), 0.5 * (hi - lo)


def normalise_inputs(u0, theta, theta_b, norm=None):
    """Network inputs; norm: the 10-float normalisation (default: the box's)."""
    if norm is None:
        c, h = default_input_norm()
    else:
        c, h = np.asarray(norm[0:3], dtype=np.float64), np.asarray(norm[3:6], dtype=np.float64)
    s = np.column_stack(surrogate_coords(u0, theta, theta_b))
    return (s - c) / h


def table_lerp_f32(tab, x0, dx, x):
    """blackholes/src/device_util.cuh:74-80 in float32 arithmetic."""
    tab = np.asarray(tab, dtype=np.float32)
    n = len(tab)
    x = np.asarray(x, dtype=np.float32)
    f = (x - np.float32(x0)) / np.float32(dx)
    f = np.minimum(np.maximum(f, np.float32(0)), np.float32(n - 2) + np.float32(0.9999))
    i = f.astype(np.int32)
    f = f - i.astype(np.float32)
    return tab[i] + (tab[i + 1] - tab[i]) * f


def exact_lnf(u0, theta, theta_b, k2, k2_l0, k2_dl):
    u0 = np.asarray(u0, dtype=np.float64)
    lth = np.log(np.asarray(theta, dtype=np.float64))
    ls = np.log(np.maximum(np.sin(np.asarray(theta_b, dtype=np.float64)), 1e-3))
    lnX = u0 * np.log(10.0) - np.log(2.0 / 9.0) - 2.0 * lth - ls
    return exact_lnf_s(lnX, lth, ls, k2, k2_l0, k2_dl)


def exact_lnf_s(lnX, lth, ls, k2, k2_l0, k2_dl):
    """The same in the surrogate coordinates (ln X, ln Theta, ln sin theta_B)."""
    lnX, lth, ls = (np.asarray(v, dtype=np.float64) for v in (lnX, lth, ls))
    X, x3 = np.exp(lnX), np.exp(lnX / 3.0)
    brk = np.sqrt(X) + 1.887748625363387 * np.sqrt(x3)
    lk2 = table_lerp_f32(k2, k2_l0, k2_dl, lth.astype(np.float32)).astype(np.float64)
    fj = LN_FJ_CONST + 2.0 * lth + ls - lk2 + 2.0 * np.log(brk) - x3
    u0ln10 = lnX + np.log(2.0 / 9.0) + 2.0 * lth + ls
    fa = fj - np.log(2.0) - lth - 2.0 * u0ln10
    return fj, fa


def leung_lnf_scipy(u0, theta, theta_b):
    """ln f_j from the Leung et al. (2011) formula with K2 from scipy (no table)."""
    from scipy.special import kve
    th = np.asarray(theta, dtype=np.float64)
    sinb = np.sin(np.asarray(theta_b, dtype=np.float64))
    X = 10.0 ** np.asarray(u0, dtype=np.float64) / (2.0 / 9.0 * th * th * sinb)
    lnk2 = np.log(kve(2, 1.0 / th)) - 1.0 / th
    return (np.log(np.sqrt(2.0) * np.pi / 3.0 * 2.0 / 9.0) + 2 * np.log(th) + np.log(sinb) - lnk2
            + 2 * np.log(np.sqrt(X) + 2 ** (11 / 12) * X ** (1 / 6)) - X ** (1 / 3))
