"""Oracle: the MT19937 generator behind the reference's rotation draws.  Test infrastructure only.

corr3pcf_kernels.py:119-125 draws each rotation's three uniforms inside a Numba kernel with
``np.random.seed(s); np.random.rand(3)``.  Numba (unpinned version; its generator has been the same since
its random module was introduced) implements these as Matsumoto & Nishimura's MT19937 (ACM TOMACS 8(1),
1998; reference code mt19937ar.c): ``seed(s)`` is ``init_genrand(s)`` (Knuth's multiplier 1812433253) and
each double is ``genrand_res53`` (27 + 26 high bits of two consecutive 32-bit outputs).  This module
restates those published steps in pure Python so that the rotation draws are pinned without Numba:
against the C++ standard's known value (std::mt19937 seeded 5489 gives 4123659995 as its 10,000th
output) and against libstdc++'s std::mt19937 for other seeds.
"""

N = 624
M = 397
MATRIX_A = 0x9908B0DF
UPPER = 0x80000000
LOWER = 0x7FFFFFFF
MASK32 = 0xFFFFFFFF


class MT19937:
    def __init__(self, seed):
        self.mt = [0] * N
        self.init_genrand(seed)

    def init_genrand(self, s):
        mt = self.mt
        mt[0] = int(s) & MASK32
        for i in range(1, N):
            prev = mt[i - 1]
            mt[i] = (1812433253 * (prev ^ (prev >> 30)) + i) & MASK32
        self.mti = N

    def _twist(self):
        mt = self.mt
        for kk in range(N):
            y = (mt[kk] & UPPER) | (mt[(kk + 1) % N] & LOWER)
            mt[kk] = mt[(kk + M) % N] ^ (y >> 1) ^ (MATRIX_A if y & 1 else 0)
        self.mti = 0

    def genrand_int32(self):
        if self.mti >= N:
            self._twist()
        y = self.mt[self.mti]
        self.mti += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & MASK32

    def genrand_res53(self):
        a = self.genrand_int32() >> 5
        b = self.genrand_int32() >> 6
        return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0)


def uniforms(seed, n):
    """``np.random.seed(seed); np.random.rand(n)`` as Numba computes it."""
    g = MT19937(seed)
    return [g.genrand_res53() for _ in range(n)]
