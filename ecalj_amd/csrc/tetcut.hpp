// tetcut.hpp -- linear-tetrahedron integration of one band pair over one
// tetrahedron: the per-tetrahedron part of the chi0 histogram weights
// (tetwt5x_dtet4, S/tetwt5.f90:514-787, with tetrahedron_matrix_linear = F
// or T; the distribution function of S/m_tetrakbt.f90:441-466).
//
// For a tetrahedron T on which the occupied energy ea, the unoccupied energy
// eb and the excitation w = (eb - ea)/2 are linear, bin i of the histogram
// mesh f_0 < f_1 < ... < f_nwhis receives
//     W_i = vol{ k in T : ea(k) < E_F < eb(k),  f_{i-1} <= w(k) < f_i }.
// {ea < E_F < eb} is T clipped by two planes.  Clipping a tetrahedron by one
// plane leaves at most three tetrahedra (clip_code), so the region is at most
// nine pieces, and on each piece w is linear with the cumulative volume
//     F(t) = vol{ k in piece : w(k) <= t } / vol(piece),
// a cubic spline in t between the sorted corner values of w (the number of
// states of the linear tetrahedron method: Bloechl, Jepsen and Andersen,
// PRB 49, 16223 (1994), App. A).  W_i = vol(piece) (F(f_i) - F(f_{i-1})).
//
// Kept equal to the reference so that single bins agree to rounding:
//  * the corner values of w are pulled apart by -3e-8, -2e-8, -1e-8, 0 Ry
//    before F is formed (no denominator of F can then vanish);
//  * which tetrahedra a clipped figure is cut into, and the order of their
//    vertices, including how equal energies are ranked.  Both matter because
//    the offsets act per piece: another triangulation of the same figure
//    moves weight across a bin edge by ~1e-8 / (spread of w in the piece),
//    and with matrix_linear it moves the points that carry the bins;
//  * the corner each cut point is interpolated from (clip_vertex), so that a
//    corner on E_F yields cut points equal to it and ties rank alike.
//
// tetrahedron_matrix_linear (S/tetwt5.f90:694-709, 751-759, 788-840) splits
// 4 W_i of each piece over the corners of the whole tetrahedron T, as the
// barycentric coordinates in T of one point of the piece: the vertex average
// of the piece's cross-section w = (f_{i-1} + f_i)/2, taken at the piece's
// lowest / highest corner below / above its range (corner_weights).  The
// reference finds a piece corner's coordinates by inverting the corner
// k-vectors of T; here each vertex carries them, cut along the edges with
// the same interpolation as its position.
// Everything is plain C++ (host and device), so the same code is checked on
// the host against the oracle and runs in the tetrahedron kernel.
#pragma once

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#define TET_HD __host__ __device__
#else
#define TET_HD
#endif

// No fused multiply-adds in the interpolations and the integration: the
// cut points must round as the reference's (x86, unfused) do.  Where a
// piece has equal corner values of w (the Fermi-surface corners of a pair
// whose occupied band is flat, a core state: all at w = (E_F - e_core)/2),
// their order follows that rounding, and with matrix_linear a bin below or
// above such a piece goes wholly to its first or last corner in that order.
#if defined(__clang__)
#define TET_UNFUSED _Pragma("clang fp contract(off)")
#else
#define TET_UNFUSED
#endif

namespace ecalj {
namespace tet {

// A vertex of a piece: position in k-space, excitation w, the value that
// the next clip tests (eb - E_F for pieces of the occupied clip), and the
// barycentric coordinates in the tetrahedron's four corners.
struct Vert {
  double k[3];
  double w;
  double e2;
  double b[4];
};

// Pieces of a tetrahedron clipped to {g > 0}.  The corners carry labels
// 0..3 in ascending order of their key (the clip value up to rounding), so
// the n corners inside are the labels 4-n..3.  A piece is four 4-bit vertex
// codes a | b << 2: corner a when a == b, else the zero of g on the edge from
// the inside corner a to the outside corner b.  clip_code(n) packs the
// pieces for n inside corners, 16 bits each.
TET_HD constexpr uint64_t vcode(int a, int b) { return (uint64_t)(a | (b << 2)); }
TET_HD constexpr uint64_t piece(uint64_t v0, uint64_t v1, uint64_t v2, uint64_t v3) {
  return v0 | (v1 << 4) | (v2 << 8) | (v3 << 12);
}
TET_HD inline uint64_t clip_code(int n) {
  constexpr uint64_t one = piece(vcode(3, 3), vcode(3, 1), vcode(3, 0), vcode(3, 2));
  constexpr uint64_t two = piece(vcode(3, 3), vcode(2, 2), vcode(3, 1), vcode(3, 0)) |
                           piece(vcode(2, 2), vcode(3, 0), vcode(2, 0), vcode(3, 1)) << 16 |
                           piece(vcode(2, 2), vcode(3, 1), vcode(2, 0), vcode(2, 1)) << 32;
  constexpr uint64_t three = piece(vcode(2, 2), vcode(3, 3), vcode(2, 0), vcode(1, 1)) |
                             piece(vcode(3, 3), vcode(3, 0), vcode(1, 0), vcode(2, 0)) << 16 |
                             piece(vcode(3, 3), vcode(1, 1), vcode(1, 0), vcode(2, 0)) << 32;
  constexpr uint64_t four = piece(vcode(0, 0), vcode(1, 1), vcode(2, 2), vcode(3, 3));
  return n == 1 ? one : n == 2 ? two : n == 3 ? three : four;
}
TET_HD inline int clip_pieces(int n) { return (n == 2 || n == 3) ? 3 : (n == 1 || n == 4) ? 1 : 0; }

// label[r] = the corner of rank r when the corners are ranked by key
// ascending; equal keys rank by position, ascending (later_first = false) or
// descending (later_first = true).
TET_HD inline void rank4(const double key[4], bool later_first, int label[4]) {
  for (int c = 0; c < 4; ++c) {
    int r = 0;
    for (int d = 0; d < 4; ++d)
      if (d != c) r += key[d] < key[c] || (key[d] == key[c] && (later_first ? d > c : d < c));
    label[r] = c;
  }
}

// The point of vertex code v among the labelled corners; g is the clip value.
// The cut is interpolated from the inside corner, or from the outside one
// when it is the only corner outside (from_outside), as the reference does;
// with a corner on the clip plane (g = 0) this keeps the cut point equal to
// that corner bit for bit, and so the ranking of ties in the next clip.
TET_HD inline Vert clip_vertex(const Vert* lab, const double* g, unsigned v, bool from_outside) {
  TET_UNFUSED
  unsigned a = v & 3u, b = v >> 2;
  if (a == b) return lab[a];
  if (from_outside) {
    const unsigned x = a;
    a = b;
    b = x;
  }
  const double t = g[a] / (g[a] - g[b]);  // zero of g on the edge, from corner a
  Vert p;
  for (int d = 0; d < 3; ++d) p.k[d] = lab[a].k[d] + t * (lab[b].k[d] - lab[a].k[d]);
  p.w = lab[a].w + t * (lab[b].w - lab[a].w);
  p.e2 = lab[a].e2 + t * (lab[b].e2 - lab[a].e2);
  for (int c = 0; c < 4; ++c) p.b[c] = lab[a].b[c] + t * (lab[b].b[c] - lab[a].b[c]);
  return p;
}

// Weights of the ascending corner values e of a piece in the point that
// stands for its cross-section at ee (sum 1): the corner itself below e[0]
// or above e[3]; else the mean of the cross-section's vertices, each the
// zero of w - ee on an edge (3 vertices with one corner below ee or one
// above, 4 with two), with the corner values pulled apart as in
// cum_fraction.  At ee == e[3] exactly the reference keeps the previous
// bin's weights; the top corner is taken here.
TET_HD inline void corner_weights(const double e[4], double ee, double wcg[4]) {
  TET_UNFUSED
  const double q[4] = {e[0] - 3e-8, e[1] - 2e-8, e[2] - 1e-8, e[3]};
  for (int c = 0; c < 4; ++c) wcg[c] = 0.0;
  if (ee <= e[0]) {
    wcg[0] = 1.0;
    return;
  }
  if (!(ee < q[3])) {
    wcg[3] = 1.0;
    return;
  }
  // corners below ee: 1, 2 or 3 (ee between the pulled-apart values)
  const int nb = ee < q[1] ? 1 : ee < q[2] ? 2 : 3;
  const double inv = 1.0 / (nb * (4 - nb));
  for (int lo = 0; lo < nb; ++lo)
    for (int hi = nb; hi < 4; ++hi) {
      const double d = q[hi] - q[lo];
      wcg[lo] += (q[hi] - ee) / d * inv;
      wcg[hi] += (ee - q[lo]) / d * inv;
    }
}

// Cumulative volume fraction F(t) of a piece whose corner values of w,
// ascending and pulled apart, are s[0] < s[1] < s[2] < s[3].
TET_HD inline double cum_fraction(const double s[4], double t) {
  TET_UNFUSED
  if (t < s[0]) return 0.0;
  if (t >= s[3]) return 1.0;
  const double d10 = s[1] - s[0], d20 = s[2] - s[0], d30 = s[3] - s[0];
  if (t < s[1]) {
    const double u = t - s[0];
    return u * u * u / (d10 * d20 * d30);
  }
  if (t < s[2]) {
    const double u = t - s[1], d21 = s[2] - s[1], d31 = s[3] - s[1];
    return (d10 * d10 + 3.0 * d10 * u + 3.0 * u * u - (d20 + d31) * u * u * u / (d21 * d31)) / (d20 * d30);
  }
  const double u = s[3] - t;
  return 1.0 - u * u * u / (d30 * (s[3] - s[1]) * (s[3] - s[2]));
}

// First i in 1..n with x < f[i]; n + 1 if none.
TET_HD inline int first_edge_above(const double* f, int n, double x) {
  int lo = 1, hi = n + 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (x < f[mid]) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Histogram of w over one piece: emit(i, W_i) for the bins it covers, or
// with Linear (matrix_linear) emit(i, 4 W_i, lam), lam the coordinates in
// the tetrahedron's corners that carry the bin.  false when w leaves the
// histogram mesh.
template <bool Linear, class Emit>
TET_HD inline bool bin_piece(const Vert* p, const double* frhis, int nwhis, Emit& emit) {
  TET_UNFUSED
  double s[4];
  int ord[4];
  for (int i = 0; i < 4; ++i) {
    s[i] = p[i].w;
    ord[i] = i;
  }
  for (int i = 1; i < 4; ++i)  // ascending; equal values keep their order
    for (int j = i; j > 0 && s[j - 1] > s[j]; --j) {
      const double x = s[j - 1];
      s[j - 1] = s[j];
      s[j] = x;
      const int o = ord[j - 1];
      ord[j - 1] = ord[j];
      ord[j] = o;
    }
  const double sr[4] = {s[0], s[1], s[2], s[3]};
  const int i0 = first_edge_above(frhis, nwhis, s[0]);
  const int i1 = first_edge_above(frhis, nwhis, s[3]);
  if (i0 > nwhis || i1 > nwhis) return false;
  double e[3][3];
  for (int v = 0; v < 3; ++v)
    for (int d = 0; d < 3; ++d) e[v][d] = p[v + 1].k[d] - p[0].k[d];
  const double vol = fabs(e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                          e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                          e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0])) / 6.0;
  s[0] -= 3e-8;
  s[1] -= 2e-8;
  s[2] -= 1e-8;
  double below = 0.0;
  for (int i = i0; i <= i1; ++i) {
    const double now = frhis[i] > s[3] ? 1.0 : cum_fraction(s, frhis[i]);
    if constexpr (Linear) {
      double wcg[4], lam[4];
      corner_weights(sr, 0.5 * (frhis[i - 1] + frhis[i]), wcg);
      for (int c = 0; c < 4; ++c) {
        lam[c] = 0.0;
        for (int j = 0; j < 4; ++j) lam[c] += p[ord[j]].b[c] * wcg[j];
      }
      emit(i, 4.0 * vol * (now - below), lam);
    } else {
      emit(i, vol * (now - below));
    }
    below = now;
  }
  return true;
}

// All bins of one band pair over one tetrahedron (corner positions kc,
// energies ea, eb at the corners).  false when w leaves the histogram mesh.
template <bool Linear = false, class Emit>
TET_HD inline bool pair_histogram(const double kc[4][3], const double ea[4], const double eb[4], double ef,
                                  const double* frhis, int nwhis, Emit& emit) {
  TET_UNFUSED
  // clip 1: the occupied part, g = E_F - ea > 0; ranked by ea descending,
  // equal energies later corner first
  Vert lab[4];
  double g[4], key[4];
  int lb[4];
  for (int c = 0; c < 4; ++c) key[c] = -ea[c];
  rank4(key, true, lb);
  int nin = 0;
  for (int r = 0; r < 4; ++r) {
    const int c = lb[r];
    for (int d = 0; d < 3; ++d) lab[r].k[d] = kc[c][d];
    lab[r].w = 0.5 * (eb[c] - ea[c]);
    lab[r].e2 = eb[c] - ef;
    for (int d = 0; d < 4; ++d) lab[r].b[d] = d == c ? 1.0 : 0.0;
    g[r] = ef - ea[c];
    nin += g[r] > 0.0;
  }
  const uint64_t code1 = clip_code(nin);
  bool ok = true;
  for (int p1 = 0; p1 < clip_pieces(nin); ++p1) {
    // clip 2 of this piece: the unoccupied part, g = eb - E_F > 0; ranked by
    // eb ascending, equal energies earlier vertex first
    Vert v[4], lab2[4];
    double g2[4];
    int lb2[4];
    for (int i = 0; i < 4; ++i) v[i] = clip_vertex(lab, g, (unsigned)(code1 >> (16 * p1 + 4 * i)) & 15u, nin == 3);
    for (int i = 0; i < 4; ++i) g2[i] = v[i].e2;
    rank4(g2, false, lb2);
    int nin2 = 0;
    for (int r = 0; r < 4; ++r) {
      lab2[r] = v[lb2[r]];
      nin2 += lab2[r].e2 > 0.0;
    }
    for (int r = 0; r < 4; ++r) g2[r] = lab2[r].e2;
    const uint64_t code2 = clip_code(nin2);
    for (int p2 = 0; p2 < clip_pieces(nin2); ++p2) {
      Vert q[4];
      for (int i = 0; i < 4; ++i)
        q[i] = clip_vertex(lab2, g2, (unsigned)(code2 >> (16 * p2 + 4 * i)) & 15u, nin2 == 3);
      ok = bin_piece<Linear>(q, frhis, nwhis, emit) && ok;
    }
  }
  return ok;
}

}  // namespace tet
}  // namespace ecalj
