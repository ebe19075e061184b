"""CPU model of the dense kernel's Gray walk (superman_amd/csrc/dense_f64.hip,
dense_f64_kernel), checked against the geometry the library actually picks.

For every order n and every chunk size 2^L the planner can choose, the
model replays the kernel's schedule -- chunk-start reconstruction from
G(h 2^L), the runtime column ctz(m)+U at block starts (direction from bit
ctz(m)+1 of m -- the +A / -A plane choice -- or bit 0 of h for column L-1),
the compile-time columns ctz(j) inside a block (direction from bit c+1 of j,
or bit 0 of m for column U-1) and the term signs -- and asserts that each of
the 2^(n-1) subsets is visited exactly once, every flip goes the right way
and every term carries (-1)^|S|.  The kernel's own arithmetic is checked on
the GPU."""
import ctypes

import pytest

from superman_amd import _lib


def ctz(x):
    return (x & -x).bit_length() - 1


def replay(n, L, U):
    B = n - 1
    seen = set()

    def term(S, neg):
        assert S not in seen
        assert neg == (bin(S).count("1") & 1 == 1)
        seen.add(S)

    for h in range(1 << (B - L)):
        k0 = h << L
        S = k0 ^ (k0 >> 1)
        assert S & ((1 << max(L - 1, 0)) - 1) == 0  # reconstruction starts at column L-1
        lane_sign = -1 if h & 1 else 1
        for m in range(1 << (L - U)):
            if m:
                tz = ctz(m)
                c = tz + U
                s = (-1 if (m >> (tz + 1)) & 1 else 1) if c + 1 < L else lane_sign
                assert s == (-1 if (S >> c) & 1 else 1)
                S ^= 1 << c
            term(S, U == 0 and (m & 1) == 1)
            for j in range(1, 1 << U):
                c = ctz(j)
                s = (-1 if (j >> (c + 1)) & 1 else 1) if c + 1 < U else (-1 if m & 1 else 1)
                assert s == (-1 if (S >> c) & 1 else 1)
                S ^= 1 << c
                term(S, (j & 1) == 1)
    assert len(seen) == 1 << B


def geometry(n, waves):
    L, U = ctypes.c_int(), ctypes.c_int()
    chunks, tasks = ctypes.c_uint64(), ctypes.c_uint64()
    _lib.check(_lib.lib().spm_gray_geometry(n, waves, ctypes.byref(L), ctypes.byref(chunks), ctypes.byref(tasks),
                                            ctypes.byref(U)))
    return L.value, chunks.value, tasks.value, U.value


def unroll_bits(n):
    return geometry(n, 1)[3]


@pytest.mark.parametrize("n", range(1, 16))
def test_every_admissible_chunk_size(n):
    U = unroll_bits(n)
    lo = U + 1 if U else (1 if n >= 2 else 0)
    for L in range(lo, n):
        replay(n, L, U)
    if n == 1:
        replay(1, 0, 0)


@pytest.mark.parametrize("waves", [1, 64, 3072, 4096, 1 << 20])
def test_planner_geometry_is_admissible(waves):
    for n in range(1, 64):
        L, chunks, tasks, U = geometry(n, waves)
        assert (L >= U + 1) if U else (L >= (1 if n >= 2 else 0))
        assert L <= n - 1 and L - U <= 30
        assert chunks == 1 << (n - 1 - L)
        assert tasks == (chunks + 63) // 64
        if n <= 14:
            replay(n, L, U)
