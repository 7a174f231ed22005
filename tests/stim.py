"""Small seeded stimulus helpers for the tests (packing and random complex int16)."""
import numpy as np

from rd_scenario import pack, to_complex  # noqa: F401  (re-exported)


def rand_cx(rng, n, scale=None):
    """n packed complex int16 samples; full-range when scale is None."""
    if scale is None:
        re = rng.integers(-32768, 32768, n)
        im = rng.integers(-32768, 32768, n)
    else:
        re = np.clip(np.rint(rng.standard_normal(n) * scale / 3), -32768, 32767).astype(np.int64)
        im = np.clip(np.rint(rng.standard_normal(n) * scale / 3), -32768, 32767).astype(np.int64)
    return pack(re, im)


def rand_extreme(rng, n):
    """Samples drawn from the corner values that make sums and products saturate."""
    vals = np.array([-32768, -32767, -1, 0, 1, 32766, 32767])
    return pack(rng.choice(vals, n), rng.choice(vals, n))


def rne_int(v: int, s: int) -> int:
    q, r = divmod(v, 1 << s)
    h = 1 << (s - 1)
    if r > h or (r == h and q & 1):
        q += 1
    return max(-32768, min(32767, q))
