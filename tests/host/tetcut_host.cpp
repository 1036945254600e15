// Host build of the tetrahedron integration in ecalj_amd/csrc/tetcut.hpp,
// for comparing it with the oracle on the CPU (tests/test_tetcut_host.py).
#include "tetcut.hpp"

extern "C" int tetcut_pair_histogram(const double* kvec, const double* ea, const double* eb, double ef,
                                     const double* frhis, int nwhis, double* out) {
  double kc[4][3];
  for (int c = 0; c < 4; ++c)
    for (int d = 0; d < 3; ++d) kc[c][d] = kvec[3 * c + d];
  struct Add {
    double* o;
    void operator()(int i, double w) { o[i - 1] += w; }
  } add{out};
  return ecalj::tet::pair_histogram(kc, ea, eb, ef, frhis, nwhis, add) ? 0 : -3;
}

// tetrahedron_matrix_linear: out is nwhis x 4 (row-major), column c the
// weight carried by corner c of kvec (4 W times its share).
extern "C" int tetcut_pair_histogram_ml(const double* kvec, const double* ea, const double* eb, double ef,
                                        const double* frhis, int nwhis, double* out) {
  double kc[4][3];
  for (int c = 0; c < 4; ++c)
    for (int d = 0; d < 3; ++d) kc[c][d] = kvec[3 * c + d];
  struct Add {
    double* o;
    void operator()(int i, double w4, const double* lam) {
      for (int c = 0; c < 4; ++c) o[4 * (i - 1) + c] += w4 * lam[c];
    }
  } add{out};
  return ecalj::tet::pair_histogram<true>(kc, ea, eb, ef, frhis, nwhis, add) ? 0 : -3;
}
