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
  void skip(int64_t n) {
    for (int64_t i = 0; i < n; ++i) next();
  }
  uint64_t count() const { return count_; }

 private:
  uint32_t st_[31];
  int f_ = 3, r_ = 0;
  uint64_t count_ = 0;
};

}  // namespace dmrg
