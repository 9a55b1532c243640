"""Pins of the oracle's sample-position generators (blackholes/src/device_util.cuh:15-29,
blackholes/src/render.cu:70-80): the PCG hash, the R2 sequence and the pixel seeds.

- or_pcg is the PCG-RXS-M-XS 32-bit hash of Jarzynski & Olano, "Hash Functions for GPU Rendering"
  (JCGT 9(3), 2020): one step of the 32-bit LCG of O'Neill's PCG family (multiplier 747796405,
  increment 2891336453), then the RXS-M-XS output permutation (a random xorshift by 4 + the top
  four bits, a multiply by 277803737, an xorshift by 22). It is checked against a numpy restatement
  of that published construction, and by properties every stage has: each is a bijection of the
  32-bit words, so the hash must be one (no collisions over 2^24 inputs), and its output bits must
  be balanced and avalanche.
- or_r2_sample is Roberts' R2 sequence (2018): x_n = frac(r0 + n / phi2), y_n = frac(r1 + n / phi2^2)
  with phi2 the plastic number, the real root of x^3 = x + 1. The constants are checked against
  that root, and the sequence's star discrepancy against random points.
- or_pixel_seed draws r0, r1 as the top 24 bits of successive hashes of (p * 0x9E3779B9) ^ seed."""
import numpy as np

M32 = np.uint64(0xFFFFFFFF)


def pcg_rxs_m_xs(v):
    """Jarzynski & Olano (2020) pcg_hash: LCG step, then RXS-M-XS, in 64-bit numpy words masked to 32."""
    v = np.asarray(v, dtype=np.uint64)
    state = (v * np.uint64(747796405) + np.uint64(2891336453)) & M32
    rot = (state >> np.uint64(28)) + np.uint64(4)
    word = (((state >> rot) ^ state) * np.uint64(277803737)) & M32
    return ((word >> np.uint64(22)) ^ word).astype(np.uint32)


def test_pcg_matches_published_hash(oracle_bhm, oracle_libm):
    rng = np.random.default_rng(1)
    v = np.concatenate([np.arange(4096, dtype=np.uint32), rng.integers(0, 2 ** 32, 20000, dtype=np.uint64).astype(np.uint32),
                        np.array([0xFFFFFFFF, 0x80000000, 0x9E3779B9], dtype=np.uint32)])
    want = pcg_rxs_m_xs(v)
    for O in (oracle_bhm, oracle_libm):
        got = np.array([O.lib.or_pcg(int(x)) for x in v], dtype=np.uint32)
        assert np.array_equal(got, want)


def test_pcg_is_a_bijection_with_balanced_bits(oracle_bhm):
    v = np.arange(2 ** 24, dtype=np.uint64)
    h = pcg_rxs_m_xs(v)
    # the library agrees with the restatement on a strided subset of the same inputs
    sub = v[:: 4099]
    assert np.array_equal(np.array([oracle_bhm.lib.or_pcg(int(x)) for x in sub], dtype=np.uint32), h[:: 4099])
    assert len(np.unique(h)) == len(h)                                    # bijective stages -> no collisions
    bits = ((h[:, None] >> np.arange(32, dtype=np.uint32)) & 1).mean(axis=0)
    assert np.all(np.abs(bits - 0.5) < 1e-3)                              # each output bit balanced
    # avalanche: flipping input bit k flips each output bit about half the time
    x = np.random.default_rng(2).integers(0, 2 ** 32, 20000, dtype=np.uint64)
    hx = pcg_rxs_m_xs(x)
    for k in (0, 7, 16, 31):
        d = hx ^ pcg_rxs_m_xs(x ^ np.uint64(1 << k))
        frac = ((d[:, None] >> np.arange(32, dtype=np.uint32)) & 1).mean()
        assert abs(frac - 0.5) < 0.02


def test_r2_constants_are_the_plastic_number():
    roots = np.roots([1.0, 0.0, -1.0, -1.0])
    phi2 = float(roots[np.argmin(np.abs(roots.imag))].real)
    for _ in range(3):   # Newton polish of the companion-matrix root
        phi2 -= (phi2 ** 3 - phi2 - 1.0) / (3.0 * phi2 ** 2 - 1.0)
    assert abs(phi2 ** 3 - phi2 - 1.0) < 1e-15
    # the constants of device_util.cuh:26 as the oracle states them
    a1, a2 = 0.7548776662466927, 0.5698402909980532
    assert abs(a1 - 1.0 / phi2) < 2e-16 and abs(a2 - 1.0 / phi2 ** 2) < 2e-16


def test_r2_sequence(oracle_bhm):
    import ctypes as C
    O = oracle_bhm
    phi2 = 1.3247179572447460
    x, y = C.c_float(), C.c_float()
    pts = []
    for r0, r1 in ((0.0, 0.0), (0.3125, 0.8125)):
        for n in range(4096):
            O.lib.or_r2_sample(n, r0, r1, C.byref(x), C.byref(y))
            ex = (r0 + (n / phi2) % 1.0) % 1.0
            ey = (r1 + (n / phi2 ** 2) % 1.0) % 1.0
            # float arithmetic: within a few float ulps of the exact sequence (mod 1)
            dx = min(abs(x.value - ex), 1 - abs(x.value - ex))
            dy = min(abs(y.value - ey), 1 - abs(y.value - ey))
            assert dx < 4e-7 and dy < 4e-7, (n, x.value, ex, y.value, ey)
            assert 0.0 <= x.value < 1.0 and 0.0 <= y.value < 1.0
            if r0 == 0.0:
                pts.append((x.value, y.value))
    # low discrepancy: the fraction of the first N points in [0, a) x [0, b) tracks a * b far better
    # than random points do
    P = np.array(pts)
    grid = np.linspace(0.05, 1.0, 20)
    worst = max(abs(np.mean((P[:, 0] < a) & (P[:, 1] < b)) - a * b) for a in grid for b in grid)
    assert worst < 3e-3   # random points: ~ 1 / sqrt(4096) = 1.6e-2


def test_pixel_seed(oracle_bhm):
    import ctypes as C
    O = oracle_bhm
    r0, r1 = C.c_float(), C.c_float()
    seed = 0x9E3779B9 * 3 & 0xFFFFFFFF
    for p in (0, 1, 1023, 524288, 1048575):
        O.lib.or_pixel_seed(seed, p, C.byref(r0), C.byref(r1))
        h1 = pcg_rxs_m_xs(pcg_rxs_m_xs(np.uint32((p * 0x9E3779B9 & 0xFFFFFFFF) ^ seed)))
        h2 = pcg_rxs_m_xs(h1)
        assert r0.value == float(int(h1) >> 8) / 16777216.0
        assert r1.value == float(int(h2) >> 8) / 16777216.0
