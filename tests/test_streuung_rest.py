"""The remainder by multiplication that the device's self-collision hash uses
(gc_selbst.h: sk_rest_m, sk_rest; Lemire, Kaser and Kurz 2019): with
m = ceil(2^64 / n) taken mod 2^64, u mod n equals the high 64 bits of
(m u mod 2^64) n for every 32-bit u and n >= 1. Checked here in exact integer
arithmetic against Python's %, on the bucket counts the hash uses
(max(2N, 64), selbstkollision.rs:282), the edge values and random pairs, so the
bucket of every cell is the reference's (selbstkollision.rs:274-278)."""
import numpy as np

M64 = 1 << 64


def rest_m(n: int) -> int:
    return ((M64 - 1) // n + 1) % M64


def rest(u: int, m: int, n: int) -> int:
    return (((m * u) % M64) * n) >> 64


def test_rest_gleich_modulo():
    rng = np.random.default_rng(20261204)
    teiler = [1, 2, 3, 63, 64, 65, 127, 2 * 20023, 2 * 40758, 2 * 77465, 2 * 115065, 2 * 249592, (1 << 31) - 1,
              1 << 31, (1 << 32) - 1] + [int(x) for x in rng.integers(1, 1 << 32, 200)]
    zahlen = [0, 1, 2, (1 << 31) - 1, 1 << 31, (1 << 32) - 2, (1 << 32) - 1] + [int(x) for x in rng.integers(0, 1 << 32, 2000)]
    for n in teiler:
        m = rest_m(n)
        for u in zahlen:
            assert rest(u, m, n) == u % n, (u, n)


def test_streuung_wie_referenz():
    """The hash of a few cells, remainder by multiplication against %."""
    def roh(i, j, k):
        return ((i * 0x8DA6B343) ^ (j * 0xD8163841) ^ (k * 0xCB1AB31F)) % (1 << 32)
    for n in (64, 81516, 230130):
        m = rest_m(n)
        for i, j, k in [(0, 0, 0), (-1, 5, 7), (123, -456, 789), (-(1 << 20), 1 << 20, 3)]:
            u = roh(i % (1 << 32), j % (1 << 32), k % (1 << 32))
            assert rest(u, m, n) == u % n
