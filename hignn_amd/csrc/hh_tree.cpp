// Cluster tree, node boxes and close/far block lists (host side of
// HignnModel::Update, src/HignnModel.cpp:457-461).
//
// The tree is the reference's PCA bisection (src/HignnModel.cpp:119-198,
// src/Build.cpp:17-265) with one rank, built level by level: every node of a
// level is divided in parallel, children are numbered breadth first in the
// order of their parents, which is the numbering the reference's FIFO queue
// produces.  Where the reference leaves the result open (Eigen's SVD sign and
// rounding, std::sort's tie order, a float mean whose value depends on the
// summation order) the divide step fixes:
//   * mean and covariance in FP64, summed over the node in map order in
//     chunks of 256 (each chunk sequentially, then the chunk totals
//     sequentially);
//   * the principal axis by cyclic Jacobi rotations of the 3x3 covariance
//     (at most 50 sweeps), the eigenvector of the largest eigenvalue, signed
//     so that its largest-magnitude component is positive;
//   * projections t = (u0 d0 + u1 d1) + u2 d2 in FP64, particles ordered by
//     (t, current position), split = start + #{t <= 0};
//   * a split that would leave one side empty is not taken.
// The file is compiled with -ffp-contract=off so every product above is
// rounded on its own.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <thread>

#include "hh_internal.h"

namespace hh {
namespace {

template <class F>
void run_parallel(int64_t n, int nthreads, F &&f) {
  if (n <= 0) return;
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, n));
  if (nt == 1) {
    for (int64_t i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<int64_t> next(0);
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (int t = 0; t < nt; t++)
    pool.emplace_back([&]() {
      for (int64_t i; (i = next.fetch_add(1, std::memory_order_relaxed)) < n;) f(i);
    });
  for (auto &th : pool) th.join();
}

constexpr int64_t kChunk = 256;

// Sum of g(p) for p in [s, e): chunks of kChunk summed in order.
template <class G>
inline double ordered_sum(int64_t s, int64_t e, G &&g) {
  double acc = 0.0;
  int64_t p = s;
  while (p < e) {
    const int64_t stop = std::min(e, p + kChunk);
    double c = 0.0;
    for (; p < stop; ++p) c += g(p);
    acc += c;
  }
  return acc;
}

// Largest-eigenvalue eigenvector of the symmetric matrix
// [[xx,xy,zx],[xy,yy,yz],[zx,yz,zz]] by cyclic Jacobi.
void dominant_axis(double xx, double yy, double zz, double xy, double yz, double zx,
                   double axis[3]) {
  double A[3][3] = {{xx, xy, zx}, {xy, yy, yz}, {zx, yz, zz}};
  double V[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int sweep = 0; sweep < 50; ++sweep) {
    if (A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2] == 0.0) break;
    for (const auto &pq : pairs) {
      const int p = pq[0], q = pq[1];
      const double apq = A[p][q];
      if (apq == 0.0) continue;
      const double th = (A[q][q] - A[p][p]) / (2.0 * apq);
      const double t = (th >= 0.0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
      const double c = 1.0 / std::sqrt(t * t + 1.0);
      const double s = t * c;
      for (int k = 0; k < 3; ++k) {
        const double x = A[k][p], y = A[k][q];
        A[k][p] = c * x - s * y;
        A[k][q] = s * x + c * y;
      }
      for (int k = 0; k < 3; ++k) {
        const double x = A[p][k], y = A[q][k];
        A[p][k] = c * x - s * y;
        A[q][k] = s * x + c * y;
      }
      for (int k = 0; k < 3; ++k) {
        const double x = V[k][p], y = V[k][q];
        V[k][p] = c * x - s * y;
        V[k][q] = s * x + c * y;
      }
    }
  }
  int col = 0;
  if (A[1][1] > A[col][col]) col = 1;
  if (A[2][2] > A[col][col]) col = 2;
  int big = 0;
  for (int d = 1; d < 3; ++d)
    if (std::fabs(V[d][col]) > std::fabs(V[big][col])) big = d;
  const double sg = V[big][col] < 0.0 ? -1.0 : 1.0;
  for (int d = 0; d < 3; ++d) axis[d] = sg * V[d][col];
}

// Divides perm[s,e); returns the split position or s when not divided.
int64_t split_node(const float *x, int64_t *perm, int64_t s, int64_t e) {
  const int64_t len = e - s;
  const double inv_len = (double)len;
  double mu[3];
  for (int d = 0; d < 3; ++d)
    mu[d] = ordered_sum(s, e, [&](int64_t p) { return (double)x[3 * perm[p] + d]; }) / inv_len;
  auto c = [&](int64_t p, int d) { return (double)x[3 * perm[p] + d] - mu[d]; };
  const double xx = ordered_sum(s, e, [&](int64_t p) { return c(p, 0) * c(p, 0); });
  const double yy = ordered_sum(s, e, [&](int64_t p) { return c(p, 1) * c(p, 1); });
  const double zz = ordered_sum(s, e, [&](int64_t p) { return c(p, 2) * c(p, 2); });
  const double xy = ordered_sum(s, e, [&](int64_t p) { return c(p, 0) * c(p, 1); });
  const double yz = ordered_sum(s, e, [&](int64_t p) { return c(p, 1) * c(p, 2); });
  const double zx = ordered_sum(s, e, [&](int64_t p) { return c(p, 2) * c(p, 0); });
  double ax[3];
  dominant_axis(xx, yy, zz, xy, yz, zx, ax);
  struct Key {
    double t;
    int64_t pos;
  };
  std::vector<Key> keys((size_t)len);
  int64_t nonpos = 0;
  for (int64_t p = s; p < e; ++p) {
    const double t = (ax[0] * c(p, 0) + ax[1] * c(p, 1)) + ax[2] * c(p, 2);
    keys[(size_t)(p - s)] = {t, p};
    nonpos += (t <= 0.0);
  }
  std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
    return a.t < b.t || (a.t == b.t && a.pos < b.pos);
  });
  std::vector<int64_t> old(perm + s, perm + e);
  for (int64_t k = 0; k < len; ++k) perm[s + k] = old[(size_t)(keys[(size_t)k].pos - s)];
  const int64_t split = s + nonpos;
  return (split == s || split == e) ? s : split;
}

void build_tree(const float *x, int64_t n, int leaf, int nthreads, TreeLists &out) {
  out.perm.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) out.perm[(size_t)i] = i;
  int64_t *perm = out.perm.data();
  // initial stage (src/Build.cpp:38-97): root divided once, rows 1 and 2
  uint64_t r1[5] = {0, 0, 0, 0, 0}, r2[5] = {0, 0, 0, 0, 0};
  if (n > leaf) {
    const int64_t d = split_node(x, perm, 0, n);
    if (d != 0) {
      r1[2] = 0, r1[3] = (uint64_t)d, r1[4] = 1;
      r2[2] = (uint64_t)d, r2[3] = (uint64_t)n, r2[4] = 1;
    }
  }
  // local stage (src/Build.cpp:111-180): breadth first from the root again
  struct Node {
    int64_t s, e, lvl, c0, c1;
  };
  std::vector<Node> nodes{{0, n, 0, 0, 0}};
  std::vector<int64_t> level{0};
  while (!level.empty()) {
    std::vector<int64_t> split((size_t)level.size(), 0);
    run_parallel((int64_t)level.size(), nthreads, [&](int64_t k) {
      const Node &nd = nodes[(size_t)level[(size_t)k]];
      split[(size_t)k] = (nd.e - nd.s > leaf) ? split_node(x, perm, nd.s, nd.e) : nd.s;
    });
    std::vector<int64_t> next;
    for (size_t k = 0; k < level.size(); ++k) {
      const int64_t id = level[k];
      const int64_t sp = split[k];
      if (sp == nodes[(size_t)id].s) continue;
      const Node par = nodes[(size_t)id];
      const int64_t a = (int64_t)nodes.size(), b = a + 1;
      nodes[(size_t)id].c0 = a;
      nodes[(size_t)id].c1 = b;
      nodes.push_back({par.s, sp, par.lvl + 1, 0, 0});
      nodes.push_back({sp, par.e, par.lvl + 1, 0, 0});
      next.push_back(a);
      next.push_back(b);
    }
    level.swap(next);
  }
  // merge (src/Build.cpp:199-247): local id l >= 1 -> global l + 2
  const int64_t nloc = (int64_t)nodes.size() - 1;
  out.T = 3 + nloc;
  out.rows.assign((size_t)(5 * out.T), 0);
  uint64_t *R = out.rows.data();
  R[2] = 0, R[3] = (uint64_t)n, R[4] = 0;
  for (int c = 0; c < 5; ++c) R[5 + c] = r1[c], R[10 + c] = r2[c];
  if (nloc > 0) R[0] = 3, R[1] = 4;
  for (int64_t l = 1; l <= nloc; ++l) {
    const Node &nd = nodes[(size_t)l];
    uint64_t *row = R + 5 * (l + 2);
    if (nd.c0) row[0] = (uint64_t)(nd.c0 + 2), row[1] = (uint64_t)(nd.c1 + 2);
    row[2] = (uint64_t)nd.s, row[3] = (uint64_t)nd.e, row[4] = (uint64_t)nd.lvl;
  }
}

// Boxes of every reachable node: leaves from their particles
// (src/HignnModel.cpp:76-95), internal nodes from their children.
void compute_boxes(const float *x, TreeLists &tl, int nthreads) {
  tl.aux.assign((size_t)(6 * tl.T), 0.0f);
  // reachable nodes in breadth-first (= id) order
  std::vector<int64_t> order{0};
  for (size_t k = 0; k < order.size(); ++k) {
    const int64_t nd = order[k];
    if (!tl.leaf(nd)) {
      order.push_back((int64_t)tl.rows[5 * nd]);
      order.push_back((int64_t)tl.rows[5 * nd + 1]);
    }
  }
  std::vector<int64_t> lv;
  for (int64_t nd : order)
    if (tl.leaf(nd)) lv.push_back(nd);
  run_parallel((int64_t)lv.size(), nthreads, [&](int64_t k) {
    const int64_t nd = lv[(size_t)k];
    float b[6];
    for (int d = 0; d < 3; ++d) {
      b[2 * d] = std::numeric_limits<float>::max();
      b[2 * d + 1] = -std::numeric_limits<float>::max();
    }
    for (uint64_t p = tl.rows[5 * nd + 2]; p < tl.rows[5 * nd + 3]; ++p) {
      const float *xp = x + 3 * tl.perm[p];
      for (int d = 0; d < 3; ++d) {
        b[2 * d] = (b[2 * d] > xp[d]) ? xp[d] : b[2 * d];
        b[2 * d + 1] = (b[2 * d + 1] < xp[d]) ? xp[d] : b[2 * d + 1];
      }
    }
    std::copy(b, b + 6, &tl.aux[(size_t)(6 * nd)]);
  });
  for (auto it = order.rbegin(); it != order.rend(); ++it) {
    const int64_t nd = *it;
    if (tl.leaf(nd)) continue;
    const float *l = &tl.aux[(size_t)(6 * tl.rows[5 * nd])];
    const float *r = &tl.aux[(size_t)(6 * tl.rows[5 * nd + 1])];
    float *o = &tl.aux[(size_t)(6 * nd)];
    for (int d = 0; d < 3; ++d) {
      o[2 * d] = std::min(l[2 * d], r[2 * d]);
      o[2 * d + 1] = std::max(l[2 * d + 1], r[2 * d + 1]);
    }
  }
}

// Admissibility (src/HignnModel.cpp:482-518), float arithmetic in the
// reference's order.
bool admissible(const float *aux, int64_t p, int64_t q) {
  const float *a = aux + 6 * p, *b = aux + 6 * q;
  const bool apart = a[0] > b[1] || a[1] < b[0] || a[2] > b[3] || a[3] < b[2] ||
                     a[4] > b[5] || a[5] < b[4];
  if (!apart) return false;
  float da = 0.0f, db = 0.0f, cc = 0.0f;
  for (int j = 0; j < 3; ++j) {
    float t = a[2 * j] - a[2 * j + 1];
    da += t * t;
    t = b[2 * j] - b[2 * j + 1];
    db += t * t;
    t = a[2 * j] + a[2 * j + 1] - b[2 * j] - b[2 * j + 1];
    cc += t * t;
  }
  cc *= 0.25f;
  return cc > da && cc > db;
}

// Dual traversal (src/CloseFarCheck.cpp:181-454), nodes visited in id order
// (the breadth-first order of the reference's queue); the nodes of one level
// are independent and are processed in parallel.
void traverse(TreeLists &tl, bool sym, int64_t R, int nthreads) {
  const int64_t T = tl.T;
  std::vector<std::vector<int64_t>> cand((size_t)T), farm((size_t)T);
  cand[0] = {0};
  const float *aux = tl.aux.data();
  auto kids = [&](int64_t nd) {
    return std::pair<int64_t, int64_t>((int64_t)tl.rows[5 * nd], (int64_t)tl.rows[5 * nd + 1]);
  };
  // descendants of m, right child first (LIFO stack), filtered by pred/stop
  auto leaf_descendants = [&](int64_t m, std::vector<int64_t> &outv) {
    std::vector<int64_t> st{m};
    while (!st.empty()) {
      const int64_t c = st.back();
      st.pop_back();
      if (tl.leaf(c)) {
        outv.push_back(c);
      } else {
        st.push_back(kids(c).first);
        st.push_back(kids(c).second);
      }
    }
  };
  // (a leaf is taken even when it is not smaller than R: the reference would
  // push its child ids 0,0 -- the root -- and never terminate)
  auto bounded_descendants = [&](int64_t m, std::vector<int64_t> &outv) {
    std::vector<int64_t> st{m};
    while (!st.empty()) {
      const int64_t c = st.back();
      st.pop_back();
      if (tl.size(c) < R || tl.leaf(c)) {
        outv.push_back(c);
      } else {
        st.push_back(kids(c).first);
        st.push_back(kids(c).second);
      }
    }
  };
  std::vector<int64_t> level{0};
  std::vector<uint8_t> isleaf_level;
  while (!level.empty()) {
    run_parallel((int64_t)level.size(), nthreads, [&](int64_t k) {
      const int64_t nd = level[(size_t)k];
      std::vector<int64_t> keep;
      if (!tl.leaf(nd)) {
        for (int64_t m : cand[(size_t)nd]) {
          if (admissible(aux, nd, m) && tl.size(nd) < R && tl.size(m) < R) {
            farm[(size_t)nd].push_back(m);
          } else if (!tl.leaf(m)) {
            keep.push_back(kids(m).first);
            keep.push_back(kids(m).second);
          } else {
            keep.push_back(m);
          }
        }
        cand[(size_t)kids(nd).first] = keep;
        cand[(size_t)kids(nd).second] = std::move(keep);
      } else {
        for (int64_t m : cand[(size_t)nd]) {
          if (admissible(aux, nd, m)) {
            if (tl.size(m) < R)
              farm[(size_t)nd].push_back(m);
            else
              bounded_descendants(m, farm[(size_t)nd]);
          } else if (tl.leaf(m)) {
            keep.push_back(m);
          } else {
            leaf_descendants(m, keep);
          }
        }
        cand[(size_t)nd] = std::move(keep);
      }
    });
    std::vector<int64_t> next;
    for (int64_t nd : level) {
      if (tl.leaf(nd)) {
        tl.leaves.push_back(nd);
      } else {
        next.push_back(kids(nd).first);
        next.push_back(kids(nd).second);
      }
    }
    level.swap(next);
  }
  for (int64_t a : tl.leaves)
    for (int64_t b : cand[(size_t)a])
      if (!sym || b >= a) tl.closeI.push_back(a), tl.closeJ.push_back(b);
  for (int64_t a = 0; a < T; ++a)
    for (int64_t b : farm[(size_t)a])
      if (!sym || b >= a) tl.farI.push_back(a), tl.farJ.push_back(b);
}

}  // namespace

void tree_principal_axis(const double cov[6], double axis[3]) {
  dominant_axis(cov[0], cov[1], cov[2], cov[3], cov[4], cov[5], axis);
}

void build_tree_lists(const float *x, int64_t n, int leaf, bool use_symmetry,
                      int64_t max_rel_coord, int nthreads, TreeLists &out) {
  out = TreeLists();
  out.n = n;
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  build_tree(x, n, leaf, nthreads, out);
  compute_boxes(x, out, nthreads);
  traverse(out, use_symmetry, max_rel_coord, nthreads);
}

}  // namespace hh
