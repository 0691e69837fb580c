// Compile-time comparator networks for the median test (host and device).
//
// Batcher's odd-even merge sort on NS = 2^k inputs (K. E. Batcher, "Sorting
// networks and their applications", AFIPS 1968), iterative form: for every
// merge size p and distance k, compare (i + j, i + j + k) when both lie in the
// same block of 2p.  Every comparator puts the minimum at the lower index.
//
// Two prunings, both exact:
//   * inputs at positions >= M are +inf.  A comparator (x, y), x < y, leaves
//     +inf at y and v[x] at x whenever v[y] = +inf, so by induction positions
//     >= M stay +inf and every comparator touching them is a no-op;
//   * only outputs 0..R are read.  Walking the network backwards, a
//     comparator is kept only if one of its two wires is still live.
// The kept comparators leave positions 0..R holding the R+1 smallest values
// in ascending order, and positions 0..M-1 a permutation of the inputs.
#pragma once

namespace pivx {

struct CmpNet {
  int n = 0;
  unsigned char a[1024] = {};
  unsigned char b[1024] = {};
};

constexpr CmpNet make_select_net(int NS, int M, int R) {
  CmpNet all{};
  for (int p = 1; p < NS; p <<= 1)
    for (int k = p; k >= 1; k >>= 1)
      for (int j = k % p; j + k < NS; j += 2 * k)
        for (int i = 0; i < k && i + j + k < NS; ++i)
          if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) {
            all.a[all.n] = (unsigned char)(i + j);
            all.b[all.n] = (unsigned char)(i + j + k);
            ++all.n;
          }
  bool live[256] = {};
  for (int i = 0; i <= R && i < NS; ++i) live[i] = true;
  bool keep[1024] = {};
  for (int c = all.n - 1; c >= 0; --c) {
    const int x = all.a[c], y = all.b[c];
    if (y >= M) continue;
    if (!live[x] && !live[y]) continue;
    keep[c] = true;
    live[x] = live[y] = true;
  }
  CmpNet out{};
  for (int c = 0; c < all.n; ++c)
    if (keep[c]) {
      out.a[out.n] = all.a[c];
      out.b[out.n] = all.b[c];
      ++out.n;
    }
  return out;
}

constexpr int pow2_ceil(int m) {
  int p = 1;
  while (p < m) p <<= 1;
  return p;
}

// Network selecting the order statistics 0..M/2 of M values (the median of
// any count n <= M of finite values, padded with +inf, lies in that range).
template <int M>
struct MedianNet {
  static constexpr int NS = pow2_ceil(M);
  static constexpr CmpNet net = make_select_net(NS, M, M / 2);
};

}  // namespace pivx
