// glibc-compatible rand() stream.  The reference draws every random number
// from the C library's unseeded rand() (seed 1) and fills matrices column by
// column with 0.01*(rand()%100) (src/tensor/real_matrix.cc:201-208).  This is
// the published glibc random_r TYPE_3 additive feedback generator: a 31-word
// state seeded by the Park-Miller LCG, state[f] += state[r] with the pointers
// 3 apart, output state[f] >> 1, after 310 discarded draws.  Keeping our own
// copy makes the stream independent of the host C library and of any other
// rand() caller in the process.
#pragma once

#include <cstdint>
#include <vector>

namespace dmrg {

class GlibcRand {
 public:
  explicit GlibcRand(uint32_t seed = 1) { reseed(seed); }
  void reseed(uint32_t seed) {
    if (seed == 0) seed = 1;
    int64_t word = seed;
    st_[0] = seed;
    for (int i = 1; i < 31; ++i) {
      const int64_t hi = word / 127773, lo = word % 127773;
      word = 16807 * lo - 2836 * hi;
      if (word < 0) word += 2147483647;
      st_[i] = (uint32_t)word;
    }
    f_ = 3;
    r_ = 0;
    for (int i = 0; i < 310; ++i) next();
    count_ = 0;
  }
  int next() {
    const uint32_t v = st_[f_] + st_[r_];
    st_[f_] = v;
    if (++f_ == 31) f_ = 0;
    if (++r_ == 31) r_ = 0;
    ++count_;
    return (int)(v >> 1);
  }
  // RealMatrix::RandomMatrix: row-major rows x cols, filled column-major.
  void random_matrix(int rows, int cols, double* out) {
    for (int j = 0; j < cols; ++j)
      for (int i = 0; i < rows; ++i) out[(int64_t)i * cols + j] = 0.01 * (next() % 100);
  }
  // ComplexMatrix::RandomMatrix (complex_matrix.cc:201-208): row-major
  // interleaved (re, im), filled column-major, two draws per entry from
  // Complex(0.01*(rand()%100), 0.01*(rand()%100)).  C++ leaves the order of
  // the two calls unspecified; g++ on x86-64 evaluates the second argument
  // first, so the first draw of a pair is the imaginary part.
  void complex_matrix(int rows, int cols, double* out) {
    for (int j = 0; j < cols; ++j)
      for (int i = 0; i < rows; ++i) {
        const double im = 0.01 * (next() % 100);
        const double re = 0.01 * (next() % 100);
        out[2 * ((int64_t)i * cols + j)] = re;
        out[2 * ((int64_t)i * cols + j) + 1] = im;
      }
  }
  // Advance by n draws.  Short skips step; long ones jump: the raw outputs
  // obey r[k] = r[k-3] + r[k-31] (mod 2^32), i.e. the characteristic
  // polynomial P(x) = x^31 - x^28 - 1, so with x^n = sum_j c_j x^j (mod P)
  // every later window follows from the current one:
  // r[m+n+t] = sum_j c_j r[m+j+t].
  void skip(int64_t n) {
    if (n < 4096) {
      for (int64_t i = 0; i < n; ++i) next();
      return;
    }
    // window w[0..60]: the 31 current values, oldest first, then 30 more
    uint32_t w[61];
    for (int j = 0; j < 31; ++j) w[j] = st_[(f_ + j) % 31];
    for (int j = 31; j < 61; ++j) w[j] = w[j - 3] + w[j - 31];
    uint32_t c[31];
    pow_mod(n, c);
    for (int t = 0; t < 31; ++t) {
      uint32_t v = 0;
      for (int j = 0; j < 31; ++j) v += c[j] * w[j + t];
      st_[t] = v;
    }
    f_ = 0;
    r_ = 28;
    count_ += (uint64_t)n;
  }
  uint64_t count() const { return count_; }

  // The generation window: w[0..30] the 31 latest raw values, oldest first,
  // and w[31..60] the next 30 raw values (the values skip() combines).
  void window(uint32_t* w) const {
    for (int j = 0; j < 31; ++j) w[j] = st_[(f_ + j) % 31];
    for (int j = 31; j < 61; ++j) w[j] = w[j - 3] + w[j - 31];
  }
  // Coefficients of x^(k step) mod P(x) for k = k0 .. k0 + count - 1, 31
  // each, into out; prev holds those of x^((k0 - 1) step) when k0 > 0.
  // With c = x^n mod P, sum_j c_j w[j + t] (t < 31) is the window after n
  // draws, so chunk k of a stream cut into chunks of `step` draws starts
  // from the window that these coefficients give.
  static void chunk_powers(int64_t step, int64_t k0, int64_t count, const uint32_t* prev, uint32_t* out) {
    uint32_t base[31], cur[31];
    pow_mod(step, base);
    if (k0 == 0) {
      for (int i = 0; i < 31; ++i) cur[i] = i == 0 ? 1u : 0u;
    } else {
      for (int i = 0; i < 31; ++i) cur[i] = prev[i];
      mul_mod(cur, base);
    }
    for (int64_t k = 0; k < count; ++k) {
      for (int i = 0; i < 31; ++i) out[31 * k + i] = cur[i];
      mul_mod(cur, base);
    }
  }

 private:
  // a = a * b mod P(x) on degree-30 coefficient vectors (mod 2^32)
  static void mul_mod(uint32_t* a, const uint32_t* b) {
    uint32_t p[61] = {0};
    for (int i = 0; i < 31; ++i)
      for (int j = 0; j < 31; ++j) p[i + j] += a[i] * b[j];
    for (int k = 60; k >= 31; --k) {  // x^k = x^(k-3) + x^(k-31)
      p[k - 3] += p[k];
      p[k - 31] += p[k];
      p[k] = 0;
    }
    for (int i = 0; i < 31; ++i) a[i] = p[i];
  }
  // c = x^n mod P(x)
  static void pow_mod(int64_t n, uint32_t* c) {
    uint32_t base[31] = {0};
    base[1] = 1;
    for (int i = 0; i < 31; ++i) c[i] = 0;
    c[0] = 1;
    while (n > 0) {
      if (n & 1) mul_mod(c, base);
      n >>= 1;
      if (n) mul_mod(base, base);
    }
  }

  uint32_t st_[31];
  int f_ = 3, r_ = 0;
  uint64_t count_ = 0;
};

}  // namespace dmrg
