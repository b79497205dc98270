"""glibc ``rand()`` (random_r TYPE_3 additive feedback generator), oracle copy.

The reference never seeds ``rand()`` (seed 1) and fills random matrices with
``0.01*(rand()%100)`` column by column (src/tensor/real_matrix.cc:201-208).
Algorithm: the published glibc random_r TYPE_3 generator -- a 31-word state
seeded by the Park-Miller LCG (16807, 2^31-1), front/rear pointers 3 apart,
state[f] += state[r] (mod 2^32), output state[f] >> 1, 310 warm-up draws.
"""
import numpy as np


class GlibcRand:
    def __init__(self, seed=1):
        if seed == 0:
            seed = 1
        st = [0] * 31
        st[0] = seed & 0xFFFFFFFF
        word = seed
        for i in range(1, 31):
            hi, lo = word // 127773, word % 127773
            word = 16807 * lo - 2836 * hi
            if word < 0:
                word += 2147483647
            st[i] = word & 0xFFFFFFFF
        self._st = st
        self._f = 3
        self._r = 0
        self.count = 0
        self.draws(310)
        self.count = 0

    def draws(self, n):
        """n successive rand() values as an int64 array."""
        st, f, r = self._st, self._f, self._r
        out = np.empty(n, dtype=np.int64)
        for k in range(n):
            v = (st[f] + st[r]) & 0xFFFFFFFF
            st[f] = v
            out[k] = v >> 1
            f += 1
            if f == 31:
                f = 0
            r += 1
            if r == 31:
                r = 0
        self._f, self._r = f, r
        self.count += n
        return out

    def rand(self):
        return int(self.draws(1)[0])

    def skip(self, n):
        self.draws(n)

    def random_matrix(self, rows, cols):
        """RealMatrix::RandomMatrix (real_matrix.cc:201-208): column-major fill."""
        vals = self.draws(rows * cols)
        return (0.01 * (vals % 100)).reshape(cols, rows).T.copy()

    def complex_matrix(self, rows, cols):
        """ComplexMatrix::RandomMatrix (complex_matrix.cc:201-208): column-major
        fill, two draws per entry, Complex(0.01*(rand()%100), 0.01*(rand()%100)).
        C++ leaves the order of the two calls unspecified; g++ on x86-64
        evaluates the second argument first, so the first draw of a pair is
        the imaginary part and the second the real part."""
        vals = 0.01 * (self.draws(2 * rows * cols) % 100)
        z = vals[1::2] + 1j * vals[0::2]
        return z.reshape(cols, rows).T.copy()
