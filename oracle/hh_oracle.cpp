// hh_oracle.cpp -- CPU restatement of the H-HIGNN mobility matvec u = M(x) f.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/, the smoke
// check and the CPU-baseline leg of bench.py load it as the checker.  The
// product (hignn_amd/libhh.so) never links, loads or calls it.
//
// Every function cites the reference file:line it restates.  Where the
// reference is under-determined (Eigen JacobiSVD sign and rounding, the
// unstable std::sort tie order, the order of Kokkos reductions and atomics)
// this oracle fixes a deterministic choice; the product implements the very
// same choice independently so that tree, permutation and block lists compare
// bit for bit:
//   * node mean and covariance: FP64 sums over the node's particles in map
//     order, in chunks of 256 summed sequentially, chunk totals summed
//     sequentially;
//   * principal axis: cyclic Jacobi on the 3x3 FP64 covariance (50 sweeps at
//     most), eigenvector of the largest eigenvalue, sign such that its
//     largest-magnitude component is positive (first index on ties);
//   * projection t = (u0*d0 + u1*d1) + u2*d2 in FP64, stable sort by t,
//     split = start + #{t <= 0};
//   * a split that leaves one side empty is not taken (the node stays a leaf);
//   * N <= leaf leaves the root a leaf (src/Build.cpp:239-240 would write
//     child ids 3,4 beyond a 3-row tree);
//   * ACA pivots: first index wins among equal determinants;
//   * the far-node descent below maxRelativeCoord takes a leaf even when it
//     is not smaller than the bound (src/CloseFarCheck.cpp:242-257 would
//     push the leaf's child ids 0,0 and loop).
// Build flags: -O3 -ffp-contract=off (no fused multiply-add anywhere).
//
// Parity status per function (how each is pinned) is listed in the header of
// oracle/oracle.py.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <queue>
#include <stack>
#include <thread>
#include <vector>

namespace {

constexpr int H1 = 128, H2 = 512, NOUT = 9;
constexpr int64_t W1_OFF = 0;
constexpr int64_t B1_OFF = W1_OFF + H1 * 3;
constexpr int64_t W2_OFF = B1_OFF + H1;
constexpr int64_t B2_OFF = W2_OFF + (int64_t)H2 * H1;
constexpr int64_t W3_OFF = B2_OFF + H2;
constexpr int64_t B3_OFF = W3_OFF + NOUT * H2;
constexpr int64_t W_TOTAL = B3_OFF + NOUT;  // 71,177 floats

int num_threads() {
  unsigned h = std::thread::hardware_concurrency();
  return h == 0 ? 1 : (int)h;
}

template <class F>
void parallel_for(int64_t n, F f, int nthreads = 0) {
  if (nthreads <= 0) nthreads = num_threads();
  if (n <= 0) return;
  if (nthreads == 1 || n == 1) {
    for (int64_t i = 0; i < n; i++) f(i, 0);
    return;
  }
  std::atomic<int64_t> next(0);
  std::vector<std::thread> th;
  int nt = (int)std::min<int64_t>(nthreads, n);
  for (int t = 0; t < nt; t++)
    th.emplace_back([&, t]() {
      for (;;) {
        int64_t i = next.fetch_add(1);
        if (i >= n) break;
        f(i, t);
      }
    });
  for (auto &x : th) x.join();
}

// ---------------------------------------------------------------------------
// Two-body pair function Net.forward (python/convert.py:11-20) over the
// Sequential(Linear(3,128), Tanh, Linear(128,512), Tanh, Linear(512,9)) of
// nn/two_body_unbounded.pkl.  y = W3 tanh(W2 tanh(W1 r + b1) + b2) + b3;
// y[0,4,8] += 1; y /= |r|; |r| < 1e-5 -> identity.
// Batched over P displacements; sums run over the input index in ascending
// order (the matrix layout of the batched loop does not change that order).
// ---------------------------------------------------------------------------
// Pairs are processed 16 at a time so that the loops vectorise across pairs;
// each pair's sums still accumulate strictly in ascending input order.
//
// tanh of the FP32 path: evaluated in double and rounded to float, as
// sign(x) (1 - e) / (1 + e) with e = exp(-2|x|) = 2^y, y = -2|x| log2 e,
// split into y = k + f (k = floor y, f in [0, 1)); 2^f = exp(f ln 2) by its
// Taylor series to degree 11 (remainder below (ln 2)^12 / 12! < 2.5e-11,
// far under the float rounding), 2^k assembled in the exponent bits.  Plain
// arithmetic, so the per-pair loops vectorise; libm's scalar tanhf made the
// oracle ~20x slower.  The FP64 path keeps std::tanh.
inline double exp2_nonpos(double y) {  // 2^y for y <= 0
  if (y < -1000.0) return 0.0;
  const double k = std::floor(y);
  const double z = (y - k) * 0.6931471805599453;
  double p = 1.0 / 39916800.0;  // 1/11!
  p = p * z + 1.0 / 3628800.0;
  p = p * z + 1.0 / 362880.0;
  p = p * z + 1.0 / 40320.0;
  p = p * z + 1.0 / 5040.0;
  p = p * z + 1.0 / 720.0;
  p = p * z + 1.0 / 120.0;
  p = p * z + 1.0 / 24.0;
  p = p * z + 1.0 / 6.0;
  p = p * z + 0.5;
  p = p * z + 1.0;
  p = p * z + 1.0;
  const int64_t bits = ((int64_t)k + 1023) << 52;
  double two_k;
  std::memcpy(&two_k, &bits, 8);
  return p * two_k;
}

inline float tanh_f(float x) {
  const double a = std::fabs((double)x);
  const double e = exp2_nonpos(-2.8853900817779268 * a);
  const double t = (1.0 - e) / (1.0 + e);
  return (float)(x < 0.0f ? -t : t);
}
inline float act(float x) { return tanh_f(x); }
inline double act(double x) { return std::tanh(x); }

template <class T>
void mlp_batch(const float *w, const T *r, int64_t P, T *out) {
  constexpr int PB = 16;
  const float *W1 = w + W1_OFF, *B1 = w + B1_OFF, *W2 = w + W2_OFF,
              *B2 = w + B2_OFF, *W3 = w + W3_OFF, *B3 = w + B3_OFF;
  alignas(64) T h1[H1][PB];
  alignas(64) T h2[H2][PB];
  alignas(64) T acc[PB];
  for (int64_t p0 = 0; p0 < P; p0 += PB) {
    const int np = (int)std::min<int64_t>(PB, P - p0);
    T r0[PB], r1[PB], r2[PB];
    for (int p = 0; p < PB; p++) {
      const int64_t q = p0 + (p < np ? p : 0);
      r0[p] = r[3 * q];
      r1[p] = r[3 * q + 1];
      r2[p] = r[3 * q + 2];
    }
    for (int k = 0; k < H1; k++) {
      const T b = (T)B1[k], w0 = (T)W1[3 * k], w1 = (T)W1[3 * k + 1], w2 = (T)W1[3 * k + 2];
      for (int p = 0; p < PB; p++) {
        T a = b;
        a += w0 * r0[p];
        a += w1 * r1[p];
        a += w2 * r2[p];
        h1[k][p] = act(a);
      }
    }
    for (int j = 0; j < H2; j++) {
      const float *wr = W2 + (int64_t)j * H1;
      for (int p = 0; p < PB; p++) acc[p] = (T)B2[j];
      for (int k = 0; k < H1; k++) {
        const T wk = (T)wr[k];
        for (int p = 0; p < PB; p++) acc[p] += wk * h1[k][p];
      }
      for (int p = 0; p < PB; p++) h2[j][p] = act(acc[p]);
    }
    T y[NOUT][PB];
    for (int o = 0; o < NOUT; o++) {
      const float *wr = W3 + (int64_t)o * H2;
      for (int p = 0; p < PB; p++) acc[p] = (T)B3[o];
      for (int k = 0; k < H2; k++) {
        const T wk = (T)wr[k];
        for (int p = 0; p < PB; p++) acc[p] += wk * h2[k][p];
      }
      for (int p = 0; p < PB; p++) y[o][p] = acc[p];
    }
    for (int p = 0; p < np; p++) {
      y[0][p] += (T)1;
      y[4][p] += (T)1;
      y[8][p] += (T)1;
      const T nr = std::sqrt(r0[p] * r0[p] + r1[p] * r1[p] + r2[p] * r2[p]);
      T *o9 = out + 9 * (p0 + p);
      if (nr < (T)1e-5) {
        for (int o = 0; o < 9; o++) o9[o] = (o % 4 == 0) ? (T)1 : (T)0;
      } else {
        for (int o = 0; o < 9; o++) o9[o] = y[o][p] / nr;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Cluster tree: HignnModel::Divide (src/HignnModel.cpp:119-198) with the
// deterministic choices listed at the top of the file.
// ---------------------------------------------------------------------------
struct Coords {
  const float *x;  // (n,3) original order
};

template <class F>
double chunked_sum(int64_t s, int64_t e, F f) {
  double total = 0.0;
  for (int64_t c0 = s; c0 < e; c0 += 256) {
    const int64_t c1 = std::min<int64_t>(c0 + 256, e);
    double part = 0.0;
    for (int64_t i = c0; i < c1; i++) part += f(i);
    total += part;
  }
  return total;
}

// Dominant eigenvector of a symmetric 3x3 matrix by cyclic Jacobi rotations.
void principal_axis(const double cov[6], double u[3]) {
  // cov = xx, yy, zz, xy, yz, zx
  double a[3][3] = {{cov[0], cov[3], cov[5]},
                    {cov[3], cov[1], cov[4]},
                    {cov[5], cov[4], cov[2]}};
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  static const int PQ[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int sweep = 0; sweep < 50; sweep++) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    if (off == 0.0) break;
    for (int r = 0; r < 3; r++) {
      const int p = PQ[r][0], q = PQ[r][1];
      if (a[p][q] == 0.0) continue;
      const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
      const double sgn = theta >= 0.0 ? 1.0 : -1.0;
      const double t = sgn / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
      const double c = 1.0 / std::sqrt(t * t + 1.0);
      const double s = t * c;
      for (int k = 0; k < 3; k++) {  // A <- A J (columns p,q)
        const double akp = a[k][p], akq = a[k][q];
        a[k][p] = c * akp - s * akq;
        a[k][q] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; k++) {  // A <- J^T A (rows p,q)
        const double apk = a[p][k], aqk = a[q][k];
        a[p][k] = c * apk - s * aqk;
        a[q][k] = s * apk + c * aqk;
      }
      for (int k = 0; k < 3; k++) {  // V <- V J
        const double vkp = v[k][p], vkq = v[k][q];
        v[k][p] = c * vkp - s * vkq;
        v[k][q] = s * vkp + c * vkq;
      }
    }
  }
  int best = 0;
  for (int k = 1; k < 3; k++)
    if (a[k][k] > a[best][best]) best = k;
  for (int d = 0; d < 3; d++) u[d] = v[d][best];
  int big = 0;
  for (int d = 1; d < 3; d++)
    if (std::fabs(u[d]) > std::fabs(u[big])) big = d;
  if (u[big] < 0.0)
    for (int d = 0; d < 3; d++) u[d] = -u[d];
}

// Returns the split index; returns s when the node is not divided.
int64_t divide(const float *x, std::vector<uint64_t> &map, int64_t s,
               int64_t e) {
  const int64_t L = e - s;
  double mean[3];
  for (int d = 0; d < 3; d++)
    mean[d] = chunked_sum(s, e, [&](int64_t i) { return (double)x[3 * map[i] + d]; }) /
              (double)L;
  auto dx = [&](int64_t i, int d) { return (double)x[3 * map[i] + d] - mean[d]; };
  double cov[6];
  static const int CA[6] = {0, 1, 2, 0, 1, 2}, CB[6] = {0, 1, 2, 1, 2, 0};
  for (int c = 0; c < 6; c++)
    cov[c] = chunked_sum(s, e, [&](int64_t i) { return dx(i, CA[c]) * dx(i, CB[c]); });
  double u[3];
  principal_axis(cov, u);
  std::vector<double> t(L);
  for (int64_t i = 0; i < L; i++)
    t[i] = (u[0] * dx(s + i, 0) + u[1] * dx(s + i, 1)) + u[2] * dx(s + i, 2);
  std::vector<int64_t> idx(L);
  for (int64_t i = 0; i < L; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](int64_t a, int64_t b) { return t[a] < t[b]; });
  std::vector<uint64_t> copy(map.begin() + s, map.begin() + e);
  int64_t nle = 0;
  for (int64_t i = 0; i < L; i++) {
    map[s + i] = copy[idx[i]];
    if (t[idx[i]] <= 0.0) nle++;
  }
  const int64_t split = s + nle;
  if (split == s || split == e) return s;
  return split;
}

struct Tree {
  std::vector<uint64_t> rows;  // T x 5: c0, c1, start, end, level
  std::vector<uint64_t> map;   // tree position -> original index
  int64_t T = 0;
  uint64_t &at(int64_t n, int c) { return rows[5 * n + c]; }
  uint64_t at(int64_t n, int c) const { return rows[5 * n + c]; }
  int64_t size(int64_t n) const { return (int64_t)(at(n, 3) - at(n, 2)); }
  bool leaf(int64_t n) const { return at(n, 0) == 0; }
};

// HignnModel::Build with one rank (src/Build.cpp:17-265): the root is divided
// in the initial stage (rows 1,2), divided again from that order in the local
// stage, local nodes are numbered breadth first and stored at local id + 2,
// and the root's children are rewritten to rows 3,4 (rows 1,2 unreachable).
void build_tree(const float *x, int64_t n, int leaf, Tree &tr) {
  tr.map.resize(n);
  for (int64_t i = 0; i < n; i++) tr.map[i] = i;
  uint64_t init[3][5] = {{0, 0, 0, (uint64_t)n, 0}, {0}, {0}};
  if (n > leaf) {
    const int64_t d = divide(x, tr.map, 0, n);
    if (d != 0) {
      init[0][0] = 1;
      init[0][1] = 2;
      init[1][2] = 0;
      init[1][3] = d;
      init[1][4] = 1;
      init[2][2] = d;
      init[2][3] = n;
      init[2][4] = 1;
    }
  }
  // local stage: rows indexed by local id
  std::vector<uint64_t> loc;  // 5 per node
  loc.insert(loc.end(), {init[0][0], init[0][1], 0, (uint64_t)n, 0});
  std::queue<int64_t> q;
  q.push(0);
  int64_t next = 1;
  while (!q.empty()) {
    const int64_t nd = q.front();
    q.pop();
    const int64_t s = loc[5 * nd + 2], e = loc[5 * nd + 3];
    int64_t d = s;
    if (e - s > leaf) d = divide(x, tr.map, s, e);
    if (d == s) {
      loc[5 * nd + 0] = 0;
      loc[5 * nd + 1] = 0;
    } else {
      const int64_t c0 = next++, c1 = next++;
      loc[5 * nd + 0] = c0;
      loc[5 * nd + 1] = c1;
      const uint64_t lv = loc[5 * nd + 4] + 1;
      loc.insert(loc.end(), {0, 0, (uint64_t)s, (uint64_t)d, lv});
      loc.insert(loc.end(), {0, 0, (uint64_t)d, (uint64_t)e, lv});
      q.push(c0);
      q.push(c1);
    }
  }
  const int64_t nloc = next - 1;  // local nodes other than the root
  tr.T = 3 + nloc;
  tr.rows.assign(5 * tr.T, 0);
  for (int c = 0; c < 5; c++)
    for (int r = 0; r < 3; r++) tr.rows[5 * r + c] = init[r][c];
  for (int64_t i = 0; i < nloc; i++) {
    const uint64_t *src = &loc[5 * (i + 1)];
    uint64_t *dst = &tr.rows[5 * (3 + i)];
    if (src[0] != 0) {
      dst[0] = src[0] + 2;
      dst[1] = src[1] + 2;
    }
    dst[2] = src[2];
    dst[3] = src[3];
    dst[4] = src[4];
  }
  if (nloc > 0) {
    tr.rows[0] = 3;
    tr.rows[1] = 4;
  } else {
    tr.rows[0] = 0;
    tr.rows[1] = 0;
  }
}

// Node boxes (src/HignnModel.cpp:76-95 for leaves, src/CloseFarCheck.cpp:23-175
// bottom-up min/max for internal nodes).  Unreachable rows keep zeros.
void node_boxes(const Tree &tr, const std::vector<float> &xp, std::vector<float> &aux) {
  aux.assign(6 * tr.T, 0.0f);
  std::vector<int64_t> order;  // reachable nodes, parents before children
  std::vector<int64_t> st{0};
  while (!st.empty()) {
    int64_t nd = st.back();
    st.pop_back();
    order.push_back(nd);
    if (!tr.leaf(nd)) {
      st.push_back((int64_t)tr.at(nd, 0));
      st.push_back((int64_t)tr.at(nd, 1));
    }
  }
  for (auto it = order.rbegin(); it != order.rend(); ++it) {
    const int64_t nd = *it;
    float *a = &aux[6 * nd];
    if (tr.leaf(nd)) {
      for (int d = 0; d < 3; d++) {
        a[2 * d] = std::numeric_limits<float>::max();
        a[2 * d + 1] = -std::numeric_limits<float>::max();
      }
      for (uint64_t i = tr.at(nd, 2); i < tr.at(nd, 3); i++)
        for (int d = 0; d < 3; d++) {
          const float v = xp[3 * i + d];
          a[2 * d] = (a[2 * d] > v) ? v : a[2 * d];
          a[2 * d + 1] = (a[2 * d + 1] < v) ? v : a[2 * d + 1];
        }
    } else {
      const float *l = &aux[6 * tr.at(nd, 0)], *r = &aux[6 * tr.at(nd, 1)];
      for (int d = 0; d < 3; d++) {
        a[2 * d] = std::min(l[2 * d], r[2 * d]);
        a[2 * d + 1] = std::max(l[2 * d + 1], r[2 * d + 1]);
      }
    }
  }
}

// Admissibility HignnModel::CloseFarCheck(aux,n1,n2) (src/HignnModel.cpp:482-518).
bool is_far(const float *aux, int64_t n1, int64_t n2) {
  const float *a = aux + 6 * n1, *b = aux + 6 * n2;
  if (!(a[0] > b[1] || a[1] < b[0] || a[2] > b[3] || a[3] < b[2] || a[4] > b[5] ||
        a[5] < b[4]))
    return false;
  float diam0 = 0.0f, diam1 = 0.0f, dist = 0.0f, tmp;
  for (int j = 0; j < 3; j++) {
    tmp = a[2 * j] - a[2 * j + 1];
    diam0 += tmp * tmp;
    tmp = b[2 * j] - b[2 * j + 1];
    diam1 += tmp * tmp;
    tmp = a[2 * j] + a[2 * j + 1] - b[2 * j] - b[2 * j + 1];
    dist += tmp * tmp;
  }
  dist *= 0.25f;
  return (dist > diam0) && (dist > diam1);
}

struct Lists {
  std::vector<int64_t> leaves;            // leaf list in traversal order
  std::vector<int64_t> closeI, closeJ;    // close leaf pairs
  std::vector<int64_t> farI, farJ;        // far node pairs
};

// Dual traversal HignnModel::CloseFarCheck() (src/CloseFarCheck.cpp:181-454)
// with one rank.
void close_far(const Tree &tr, const std::vector<float> &aux, bool sym, int64_t R,
               Lists &L) {
  std::vector<std::vector<int64_t>> farM(tr.T), closeM(tr.T);
  closeM[0].push_back(0);
  std::queue<int64_t> q;
  q.push(0);
  const float *ax = aux.data();
  while (!q.empty()) {
    const int64_t nd = q.front();
    q.pop();
    if (!tr.leaf(nd)) {
      std::vector<int64_t> child;
      for (int64_t m : closeM[nd]) {
        bool far = is_far(ax, nd, m);
        far = far && tr.size(nd) < R && tr.size(m) < R;
        if (far) {
          farM[nd].push_back(m);
        } else if (!tr.leaf(m)) {
          child.push_back((int64_t)tr.at(m, 0));
          child.push_back((int64_t)tr.at(m, 1));
        } else {
          child.push_back(m);
        }
      }
      closeM[tr.at(nd, 0)] = child;
      closeM[tr.at(nd, 1)] = child;
      q.push((int64_t)tr.at(nd, 0));
      q.push((int64_t)tr.at(nd, 1));
    } else {
      std::vector<int64_t> nc;
      for (int64_t m : closeM[nd]) {
        const bool far = is_far(ax, nd, m);
        if (far) {
          if (tr.size(m) < R) {
            farM[nd].push_back(m);
          } else {
            std::stack<int64_t> st;
            st.push(m);
            while (!st.empty()) {
              const int64_t c = st.top();
              st.pop();
              // a leaf is taken even when not smaller than R (the reference
              // would push its child ids 0,0, i.e. the root, forever)
              if (tr.size(c) < R || tr.leaf(c)) {
                farM[nd].push_back(c);
              } else {
                st.push((int64_t)tr.at(c, 0));
                st.push((int64_t)tr.at(c, 1));
              }
            }
          }
        } else if (tr.leaf(m)) {
          nc.push_back(m);
        } else {
          std::stack<int64_t> st;
          st.push(m);
          while (!st.empty()) {
            const int64_t c = st.top();
            st.pop();
            if (!tr.leaf(c)) {
              st.push((int64_t)tr.at(c, 0));
              st.push((int64_t)tr.at(c, 1));
            } else {
              nc.push_back(c);
            }
          }
        }
      }
      closeM[nd] = nc;
      L.leaves.push_back(nd);
    }
  }
  for (int64_t ni : L.leaves)
    for (int64_t nj : closeM[ni])
      if (!sym || nj >= ni) {
        L.closeI.push_back(ni);
        L.closeJ.push_back(nj);
      }
  for (int64_t ni = 0; ni < tr.T; ni++)
    for (int64_t nj : farM[ni])
      if (!sym || nj >= ni) {
        L.farI.push_back(ni);
        L.farJ.push_back(nj);
      }
}

struct Model {
  int64_t n = 0;
  int leaf = 0;
  std::vector<float> x;   // original order (n,3)
  std::vector<float> xp;  // tree order
  Tree tr;
  std::vector<float> aux;
  Lists lists;
  bool built = false;
  // parameters (src/HignnModel.cpp:335-386 defaults)
  double epsilon = 0.05;
  int max_iter = 100;
  bool use_symmetry = true;
  int64_t max_rel_coord = 500000;
  std::vector<int> far_rank;  // final ACA rank per far block of the last dot
  int threads = 0;
};

// s(): symmetrise a float 3x3 like the cMat/qMat copy kernels
// (src/FarDot.cpp:441-456, 763-778): diagonal kept, off-diagonal
// 0.5 * (float(a+b)).
inline void sym9(const float *y, double *o) {
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      if (r == c)
        o[3 * r + c] = y[3 * r + c];
      else
        o[3 * r + c] = 0.5 * (double)(y[3 * r + c] + y[3 * c + r]);
    }
}

inline double det9(const double *m) {
  const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6],
               h = m[7], i = m[8];
  return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

struct Accum {
  std::vector<std::vector<double>> u;  // per thread (3n)
  Accum(int nt, int64_t n) : u(nt, std::vector<double>(3 * n, 0.0)) {}
  void reduce(double *out, int64_t n) {
    for (int64_t i = 0; i < 3 * n; i++) {
      double s = 0.0;
      for (auto &v : u) s += v[i];
      out[i] += s;
    }
  }
};

// HignnModel::CloseDot (src/CloseDot.cpp:177-299): r = x[J_k]-x[I_j] (f32),
// u[I_j] += M f[J_k]; with symmetry and J > I also u[J_k] += M f[I_j]
// (same M, not transposed).  fp: tree-ordered forces (double).
void close_dot(Model &m, const float *w, const double *fp, double *up) {
  const int nt = m.threads > 0 ? m.threads : num_threads();
  Accum acc(nt, m.n);
  const int64_t nb = (int64_t)m.lists.closeI.size();
  parallel_for(
      nb,
      [&](int64_t b, int t) {
        const int64_t I = m.lists.closeI[b], J = m.lists.closeJ[b];
        const int64_t is = m.tr.at(I, 2), ie = m.tr.at(I, 3);
        const int64_t js = m.tr.at(J, 2), je = m.tr.at(J, 3);
        const int64_t nj = je - js;
        std::vector<float> r(3 * nj), y(9 * nj);
        double *u = acc.u[t].data();
        for (int64_t i = is; i < ie; i++) {
          for (int64_t k = 0; k < nj; k++)
            for (int l = 0; l < 3; l++) r[3 * k + l] = m.xp[3 * (js + k) + l] - m.xp[3 * i + l];
          mlp_batch<float>(w, r.data(), nj, y.data());
          for (int64_t k = 0; k < nj; k++) {
            const float *M = &y[9 * k];
            for (int row = 0; row < 3; row++) {
              double s = 0.0;
              for (int col = 0; col < 3; col++) s += M[3 * row + col] * fp[3 * (js + k) + col];
              u[3 * i + row] += s;
            }
            if (m.use_symmetry && J > I) {
              for (int row = 0; row < 3; row++) {
                double s = 0.0;
                for (int col = 0; col < 3; col++) s += M[3 * row + col] * fp[3 * i + col];
                u[3 * (js + k) + row] += s;
              }
            }
          }
        }
      },
      nt);
  acc.reduce(up, m.n);
}

// Block-ACA of one far block (src/FarDot.cpp:186-1110), then its apply
// (src/FarDot.cpp:1200-1388).  Returns the final rank.  Batching in the
// reference does not change the per-block arithmetic.
struct AcaOut {
  std::vector<std::vector<double>> C, Q;
  std::vector<int64_t> ipiv, jpiv;
  std::vector<double> nu2, mu2;  // per iteration, after its stop-test update
};

int aca_block(const Model &m, const float *w, int64_t I, int64_t J, const double *fp,
              double *u, AcaOut *dbg = nullptr) {
  const int64_t is = m.tr.at(I, 2), mI = m.tr.size(I);
  const int64_t js = m.tr.at(J, 2), nJ = m.tr.size(J);
  const int maxIter = m.max_iter;
  const double eps2 = m.epsilon * m.epsilon;
  std::vector<std::vector<double>> C, Q;  // C[l]: mI*9, Q[l]: nJ*9
  std::vector<int64_t> ipiv, jpiv;
  jpiv.push_back(nJ / 2);
  double mu2 = 0.0;
  std::vector<float> r(3 * std::max(mI, nJ)), y(9 * std::max(mI, nJ));
  int it = 0;
  for (;;) {
    const int k = it;
    // 1. column pass at jpiv[k] (src/FarDot.cpp:346-458), residual (:474-513)
    const int64_t jk = jpiv[k];
    for (int64_t i = 0; i < mI; i++)
      for (int l = 0; l < 3; l++)
        r[3 * i + l] = m.xp[3 * (js + jk) + l] - m.xp[3 * (is + i) + l];
    mlp_batch<float>(w, r.data(), mI, y.data());
    std::vector<double> c(9 * mI);
    for (int64_t i = 0; i < mI; i++) sym9(&y[9 * i], &c[9 * i]);
    for (int64_t i = 0; i < mI; i++)
      for (int e = 0; e < 9; e++) {
        const int row = e / 3, col = e % 3;
        double s = 0.0;
        for (int l = 0; l < k; l++)
          for (int mm = 0; mm < 3; mm++)
            s += C[l][9 * i + 3 * row + mm] * Q[l][9 * jk + 3 * mm + col];
        c[9 * i + e] -= s;
      }
    // 2. row pivot: max signed det over unused rows (:520-582)
    int64_t ik = -1;
    double best = -std::numeric_limits<double>::max();
    for (int64_t i = 0; i < mI; i++) {
      const double dv = det9(&c[9 * i]);
      bool used = false;
      for (int l = 0; l < k; l++)
        if (ipiv[l] == i) used = true;
      if (dv > best && !used) {
        best = dv;
        ik = i;
      }
    }
    if (ik < 0) ik = 0;  // all candidates NaN: no guard in the reference
    ipiv.push_back(ik);
    // inverse by adjugate (:586-611), normalise C (:614-656)
    {
      const double *p = &c[9 * ik];
      const double a = p[0], b = p[1], cc = p[2], d = p[3], e = p[4], f = p[5],
                   g = p[6], h = p[7], ii = p[8];
      const double det = a * (e * ii - f * h) - b * (d * ii - f * g) + cc * (d * h - e * g);
      double inv[9];
      inv[0] = (e * ii - f * h) / det;
      inv[1] = -(b * ii - cc * h) / det;
      inv[2] = (b * f - cc * e) / det;
      inv[3] = -(d * ii - f * g) / det;
      inv[4] = (a * ii - cc * g) / det;
      inv[5] = -(a * f - cc * d) / det;
      inv[6] = (d * h - e * g) / det;
      inv[7] = -(a * h - b * g) / det;
      inv[8] = (a * e - b * d) / det;
      std::vector<double> cn(9 * mI);
      for (int64_t i = 0; i < mI; i++)
        for (int row = 0; row < 3; row++)
          for (int col = 0; col < 3; col++) {
            double s = 0.0;
            for (int kk = 0; kk < 3; kk++) s += c[9 * i + 3 * row + kk] * inv[3 * kk + col];
            cn[9 * i + 3 * row + col] = s;
          }
      C.push_back(std::move(cn));
    }
    // 3. row pass at ipiv[k] (:667-782), residual (:797-841)
    for (int64_t j = 0; j < nJ; j++)
      for (int l = 0; l < 3; l++)
        r[3 * j + l] = -m.xp[3 * (is + ik) + l] + m.xp[3 * (js + j) + l];
    mlp_batch<float>(w, r.data(), nJ, y.data());
    std::vector<double> qv(9 * nJ);
    for (int64_t j = 0; j < nJ; j++) sym9(&y[9 * j], &qv[9 * j]);
    for (int64_t j = 0; j < nJ; j++)
      for (int e = 0; e < 9; e++) {
        const int row = e / 3, col = e % 3;
        double s = 0.0;
        for (int l = 0; l < k; l++)
          for (int mm = 0; mm < 3; mm++)
            s += C[l][9 * ik + 3 * row + mm] * Q[l][9 * j + 3 * mm + col];
        qv[9 * j + e] -= s;
      }
    Q.push_back(std::move(qv));
    const std::vector<double> &Qk = Q[k];
    const std::vector<double> &Ck = C[k];
    // 4. column pivot over j not in jpiv[0..k] (:846-903)
    {
      int64_t jn = -1;
      double bst = -std::numeric_limits<double>::max();
      for (int64_t j = 0; j < nJ; j++) {
        const double dv = det9(&Qk[9 * j]);
        bool used = false;
        for (int l = 0; l <= k; l++)
          if (jpiv[l] == j) used = true;
        if (dv > bst && !used) {
          bst = dv;
          jn = j;
        }
      }
      if (jn < 0) jn = 0;
      jpiv.push_back(jn);
    }
    // 5. norms (:919-1037) and the mu2 cross terms for l < k-1 (:1041-1078)
    double gc[6] = {0}, gq[6] = {0};
    for (int64_t i = 0; i < mI; i++) {
      const double *p = &Ck[9 * i];
      for (int kk = 0; kk < 3; kk++)
        gc[kk] += p[kk] * p[kk] + p[kk + 3] * p[kk + 3] + p[kk + 6] * p[kk + 6];
      for (int kk = 3; kk < 6; kk++) {
        const int a = kk % 3, b = (kk + 1) % 3;
        gc[kk] += p[a] * p[b] + p[a + 3] * p[b + 3] + p[a + 6] * p[b + 6];
      }
    }
    for (int64_t j = 0; j < nJ; j++) {
      const double *p = &Qk[9 * j];
      gq[0] += p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
      gq[1] += p[3] * p[3] + p[4] * p[4] + p[5] * p[5];
      gq[2] += p[6] * p[6] + p[7] * p[7] + p[8] * p[8];
      gq[3] += p[0] * p[3] + p[1] * p[4] + p[2] * p[5];
      gq[4] += p[3] * p[6] + p[4] * p[7] + p[5] * p[8];
      gq[5] += p[6] * p[0] + p[7] * p[1] + p[8] * p[2];
    }
    const double nu2 = gc[0] * gq[0] + gc[1] * gq[1] + gc[2] * gq[2] +
                       2.0 * gq[3] * gc[3] + 2.0 * gq[4] * gc[4] + 2.0 * gq[5] * gc[5];
    mu2 += nu2;
    for (int l = 0; l < k - 1; l++)
      for (int d1 = 0; d1 < 3; d1++)
        for (int d2 = 0; d2 < 3; d2++) {
          double cd = 0.0, qd = 0.0;
          for (int64_t i = 0; i < mI; i++)
            cd += Ck[9 * i + d1] * C[l][9 * i + d2] + Ck[9 * i + d1 + 3] * C[l][9 * i + d2 + 3] +
                  Ck[9 * i + d1 + 6] * C[l][9 * i + d2 + 6];
          for (int64_t j = 0; j < nJ; j++)
            qd += Qk[9 * j + 3 * d1] * Q[l][9 * j + 3 * d2] +
                  Qk[9 * j + 3 * d1 + 1] * Q[l][9 * j + 3 * d2 + 1] +
                  Qk[9 * j + 3 * d1 + 2] * Q[l][9 * j + 3 * d2 + 2];
          mu2 += 2.0 * std::fabs(cd) * std::fabs(qd);
        }
    if (dbg) {
      dbg->nu2.push_back(nu2);
      dbg->mu2.push_back(mu2);
    }
    // 6. stop test (:1099-1108)
    it = k + 1;
    if (nu2 < mu2 * eps2 || it >= maxIter || it >= std::min(mI, nJ)) break;
  }
  if (dbg) {
    dbg->C = C;
    dbg->Q = Q;
    dbg->ipiv = ipiv;
    dbg->jpiv = jpiv;
  }
  if (!u) return it;
  // 7. apply (:1200-1388)
  for (int l = 0; l < it; l++) {
    double wv[3] = {0, 0, 0};
    for (int64_t j = 0; j < nJ; j++)
      for (int row = 0; row < 3; row++) {
        double s = 0.0;
        for (int kk = 0; kk < 3; kk++) s += Q[l][9 * j + 3 * row + kk] * fp[3 * (js + j) + kk];
        wv[row] += s;
      }
    for (int64_t i = 0; i < mI; i++)
      for (int row = 0; row < 3; row++) {
        double s = 0.0;
        for (int kk = 0; kk < 3; kk++) s += C[l][9 * i + 3 * row + kk] * wv[kk];
        u[3 * (is + i) + row] += s;
      }
  }
  if (m.use_symmetry) {
    for (int l = 0; l < it; l++) {
      double wv[3] = {0, 0, 0};
      for (int64_t i = 0; i < mI; i++)
        for (int row = 0; row < 3; row++) {
          double s = 0.0;
          for (int kk = 0; kk < 3; kk++) s += C[l][9 * i + 3 * kk + row] * fp[3 * (is + i) + kk];
          wv[row] += s;
        }
      for (int64_t j = 0; j < nJ; j++)
        for (int row = 0; row < 3; row++) {
          double s = 0.0;
          for (int kk = 0; kk < 3; kk++) s += Q[l][9 * j + 3 * kk + row] * wv[kk];
          u[3 * (js + j) + row] += s;
        }
    }
  }
  return it;
}

void far_dot(Model &m, const float *w, const double *fp, double *up) {
  const int nt = m.threads > 0 ? m.threads : num_threads();
  Accum acc(nt, m.n);
  const int64_t nb = (int64_t)m.lists.farI.size();
  m.far_rank.assign(nb, 0);
  parallel_for(
      nb,
      [&](int64_t b, int t) {
        m.far_rank[b] =
            aca_block(m, w, m.lists.farI[b], m.lists.farJ[b], fp, acc.u[t].data());
      },
      nt);
  acc.reduce(up, m.n);
}

// HignnModel::DenseDot (src/DenseDot.cpp:230-246): every leaf x leaf pair,
// u[I_j] += 0.5 (M + M^T) f[J_k], float sum of the two entries.
void dense_dot(Model &m, const float *w, const double *fp, double *up) {
  const int nt = m.threads > 0 ? m.threads : num_threads();
  Accum acc(nt, m.n);
  const int64_t nl = (int64_t)m.lists.leaves.size();
  parallel_for(
      nl * nl,
      [&](int64_t b, int t) {
        const int64_t I = m.lists.leaves[b / nl], J = m.lists.leaves[b % nl];
        const int64_t is = m.tr.at(I, 2), ie = m.tr.at(I, 3);
        const int64_t js = m.tr.at(J, 2), nj = m.tr.size(J);
        std::vector<float> r(3 * nj), y(9 * nj);
        double *u = acc.u[t].data();
        for (int64_t i = is; i < ie; i++) {
          for (int64_t k = 0; k < nj; k++)
            for (int l = 0; l < 3; l++) r[3 * k + l] = m.xp[3 * (js + k) + l] - m.xp[3 * i + l];
          mlp_batch<float>(w, r.data(), nj, y.data());
          for (int64_t k = 0; k < nj; k++) {
            const float *M = &y[9 * k];
            for (int row = 0; row < 3; row++) {
              double s = 0.0;
              for (int col = 0; col < 3; col++)
                s += 0.5 * (M[3 * row + col] + M[3 * col + row]) * fp[3 * (js + k) + col];
              u[3 * i + row] += s;
            }
          }
        }
      },
      nt);
  acc.reduce(up, m.n);
}

// FP64 evaluation of the H-matrix's own block convention with exact far
// blocks, restricted to a set of target particles (tree order):
//   close (I,J) rows i in I:   M64(x_j - x_i) f_j
//   close J side (sym, J>I):   M64(x_t - x_s) f_s   (reuse, not transposed)
//   far (I,J) rows i in I:     S(M64(x_j - x_i)) f_j
//   far J side (sym):          S(M64(x_t - x_s))^T f_s
// The reference forms these sums with an FP32 pair function and an ACA of the
// far blocks; this is their FP64 limit, used to measure the approximation.
void dense64_blocks(Model &m, const float *w, const double *fp, const int64_t *tgt,
                    int64_t ntgt, double *uout) {
  // owner lookup: for each tree position, its leaf
  std::vector<int64_t> pos_leaf(m.n, -1);
  for (int64_t lf : m.lists.leaves)
    for (uint64_t i = m.tr.at(lf, 2); i < m.tr.at(lf, 3); i++) pos_leaf[i] = lf;
  // per target: enumerate all blocks whose I or J range contains it
  struct Blk {
    int64_t s0, s1;  // source range
    bool far, jside;
  };
  std::vector<std::vector<Blk>> per_node(m.tr.T);
  auto add = [&](const std::vector<int64_t> &BI, const std::vector<int64_t> &BJ, bool far) {
    for (size_t b = 0; b < BI.size(); b++) {
      const int64_t I = BI[b], J = BJ[b];
      per_node[I].push_back({(int64_t)m.tr.at(J, 2), (int64_t)m.tr.at(J, 3), far, false});
      if (m.use_symmetry && J != I)
        per_node[J].push_back({(int64_t)m.tr.at(I, 2), (int64_t)m.tr.at(I, 3), far, true});
    }
  };
  add(m.lists.closeI, m.lists.closeJ, false);
  add(m.lists.farI, m.lists.farJ, true);
  // node ancestry: targets are covered by blocks of any ancestor node
  std::vector<int64_t> parent(m.tr.T, -1);
  for (int64_t nd = 0; nd < m.tr.T; nd++)
    if (!m.tr.leaf(nd)) {
      parent[m.tr.at(nd, 0)] = nd;
      parent[m.tr.at(nd, 1)] = nd;
    }
  parallel_for(
      ntgt,
      [&](int64_t q, int) {
        const int64_t t = tgt[q];
        double acc[3] = {0, 0, 0};
        std::vector<double> r, y;
        for (int64_t nd = pos_leaf[t]; nd >= 0; nd = parent[nd]) {
          for (const Blk &b : per_node[nd]) {
            const int64_t ns = b.s1 - b.s0;
            r.resize(3 * ns);
            y.resize(9 * ns);
            for (int64_t s = 0; s < ns; s++)
              for (int l = 0; l < 3; l++) {
                const double xs = m.xp[3 * (b.s0 + s) + l], xt = m.xp[3 * t + l];
                r[3 * s + l] = b.jside ? (xt - xs) : (xs - xt);
              }
            mlp_batch<double>(w, r.data(), ns, y.data());
            for (int64_t s = 0; s < ns; s++) {
              const double *M = &y[9 * s];
              const double *f = &fp[3 * (b.s0 + s)];
              for (int row = 0; row < 3; row++)
                for (int col = 0; col < 3; col++) {
                  double a = M[3 * row + col];
                  if (b.far) a = (row == col) ? a : 0.5 * (M[3 * row + col] + M[3 * col + row]);
                  acc[row] += a * f[col];
                }
            }
          }
        }
        for (int d = 0; d < 3; d++) uout[3 * q + d] = acc[d];
      },
      m.threads);
}

}  // namespace

// ---------------------------------------------------------------------------
// C interface used by oracle/oracle.py (ctypes).
// ---------------------------------------------------------------------------
extern "C" {

int64_t orc_weight_count() { return W_TOTAL; }

void orc_mlp_f32(const float *w, const float *r, int64_t P, float *out) {
  parallel_for((P + 1023) / 1024, [&](int64_t c, int) {
    const int64_t p0 = c * 1024, p1 = std::min<int64_t>(P, p0 + 1024);
    mlp_batch<float>(w, r + 3 * p0, p1 - p0, out + 9 * p0);
  });
}

void orc_mlp_f64(const float *w, const double *r, int64_t P, double *out) {
  parallel_for((P + 1023) / 1024, [&](int64_t c, int) {
    const int64_t p0 = c * 1024, p1 = std::min<int64_t>(P, p0 + 1024);
    mlp_batch<double>(w, r + 3 * p0, p1 - p0, out + 9 * p0);
  });
}

void orc_principal_axis(const double *cov6, double *u3) { principal_axis(cov6, u3); }

void *orc_model_new(const float *x, int64_t n, int leaf) {
  Model *m = new Model();
  m->n = n;
  m->leaf = leaf;
  m->x.assign(x, x + 3 * n);
  return m;
}

void orc_model_free(void *h) { delete (Model *)h; }

void orc_model_set_params(void *h, double epsilon, int max_iter, int use_symmetry,
                          int64_t max_rel_coord, int threads) {
  Model *m = (Model *)h;
  m->epsilon = epsilon;
  m->max_iter = max_iter;
  m->use_symmetry = use_symmetry != 0;
  m->max_rel_coord = max_rel_coord;
  m->threads = threads;
}

// HignnModel::UpdateCoord + Update (src/HignnModel.cpp:744-762, 457-461).
int orc_model_update(void *h, const float *x) {
  Model *m = (Model *)h;
  if (x) m->x.assign(x, x + 3 * m->n);
  m->tr = Tree();
  build_tree(m->x.data(), m->n, m->leaf, m->tr);
  m->xp.resize(3 * m->n);
  for (int64_t i = 0; i < m->n; i++)
    for (int d = 0; d < 3; d++) m->xp[3 * i + d] = m->x[3 * m->tr.map[i] + d];
  node_boxes(m->tr, m->xp, m->aux);
  m->lists = Lists();
  close_far(m->tr, m->aux, m->use_symmetry, m->max_rel_coord, m->lists);
  m->built = true;
  return 0;
}

// what: 0 tree rows, 1 close pairs, 2 far pairs, 3 leaves
int64_t orc_model_count(void *h, int what) {
  Model *m = (Model *)h;
  switch (what) {
    case 0: return m->tr.T;
    case 1: return (int64_t)m->lists.closeI.size();
    case 2: return (int64_t)m->lists.farI.size();
    case 3: return (int64_t)m->lists.leaves.size();
  }
  return -1;
}

void orc_model_export(void *h, uint64_t *tree, uint64_t *map, float *aux, int64_t *close_ij,
                      int64_t *far_ij, int64_t *leaves) {
  Model *m = (Model *)h;
  if (tree) std::memcpy(tree, m->tr.rows.data(), sizeof(uint64_t) * 5 * m->tr.T);
  if (map) std::memcpy(map, m->tr.map.data(), sizeof(uint64_t) * m->n);
  if (aux) std::memcpy(aux, m->aux.data(), sizeof(float) * 6 * m->tr.T);
  if (close_ij)
    for (size_t b = 0; b < m->lists.closeI.size(); b++) {
      close_ij[2 * b] = m->lists.closeI[b];
      close_ij[2 * b + 1] = m->lists.closeJ[b];
    }
  if (far_ij)
    for (size_t b = 0; b < m->lists.farI.size(); b++) {
      far_ij[2 * b] = m->lists.farI[b];
      far_ij[2 * b + 1] = m->lists.farJ[b];
    }
  if (leaves) std::memcpy(leaves, m->lists.leaves.data(), sizeof(int64_t) * m->lists.leaves.size());
}

// is_far on the model's boxes (for brute-force tests).
int orc_model_is_far(void *h, int64_t a, int64_t b) {
  Model *m = (Model *)h;
  return is_far(m->aux.data(), a, b) ? 1 : 0;
}

// HignnModel::Dot (src/HignnModel.cpp:545-634): f -> double, permute, CloseDot
// + FarDot, un-permute.  parts: 1 close only, 2 far only, 3 both.
// u_out is double (n,3) in the original order.
int orc_model_dot(void *h, const float *w, const float *f, double *u_out, int parts) {
  Model *m = (Model *)h;
  if (!m->built) orc_model_update(h, nullptr);
  std::vector<double> fp(3 * m->n), up(3 * m->n, 0.0);
  for (int64_t i = 0; i < m->n; i++)
    for (int d = 0; d < 3; d++) fp[3 * i + d] = f[3 * m->tr.map[i] + d];
  if (parts & 1) close_dot(*m, w, fp.data(), up.data());
  if (parts & 2) far_dot(*m, w, fp.data(), up.data());
  for (int64_t i = 0; i < m->n; i++)
    for (int d = 0; d < 3; d++) u_out[3 * m->tr.map[i] + d] = up[3 * i + d];
  return 0;
}

void orc_model_far_ranks(void *h, int32_t *ranks) {
  Model *m = (Model *)h;
  for (size_t b = 0; b < m->far_rank.size(); b++) ranks[b] = m->far_rank[b];
}

// HignnModel::DenseDot(u,f) (src/HignnModel.cpp:660-728, src/DenseDot.cpp).
int orc_model_dense_dot(void *h, const float *w, const float *f, double *u_out) {
  Model *m = (Model *)h;
  if (!m->built) orc_model_update(h, nullptr);
  std::vector<double> fp(3 * m->n), up(3 * m->n, 0.0);
  for (int64_t i = 0; i < m->n; i++)
    for (int d = 0; d < 3; d++) fp[3 * i + d] = f[3 * m->tr.map[i] + d];
  dense_dot(*m, w, fp.data(), up.data());
  for (int64_t i = 0; i < m->n; i++)
    for (int d = 0; d < 3; d++) u_out[3 * m->tr.map[i] + d] = up[3 * i + d];
  return 0;
}

// FP64 block-convention dense matvec at selected targets (original indices).
int orc_model_dense64_blocks(void *h, const float *w, const float *f, const int64_t *targets,
                             int64_t nt, double *u_out) {
  Model *m = (Model *)h;
  if (!m->built) orc_model_update(h, nullptr);
  std::vector<double> fp(3 * m->n);
  for (int64_t i = 0; i < m->n; i++)
    for (int d = 0; d < 3; d++) fp[3 * i + d] = f[3 * m->tr.map[i] + d];
  std::vector<int64_t> inv(m->n);
  for (int64_t i = 0; i < m->n; i++) inv[m->tr.map[i]] = i;
  std::vector<int64_t> tp(nt);
  for (int64_t q = 0; q < nt; q++) tp[q] = inv[targets[q]];
  dense64_blocks(*m, w, fp.data(), tp.data(), nt, u_out);
  return 0;
}

// FP64 ordered-pair dense matvec u_t = sum_s 0.5(M+M^T)(x_s - x_t) f_s at
// selected targets (src/DenseDot.cpp:230-246 semantics in FP64).
void orc_dense64_ordered(const float *w, const float *x, const float *f, int64_t n,
                         const int64_t *targets, int64_t nt, double *u_out) {
  parallel_for(nt, [&](int64_t q, int) {
    const int64_t t = targets[q];
    std::vector<double> r(3 * 1024), y(9 * 1024);
    double acc[3] = {0, 0, 0};
    for (int64_t s0 = 0; s0 < n; s0 += 1024) {
      const int64_t ns = std::min<int64_t>(1024, n - s0);
      for (int64_t s = 0; s < ns; s++)
        for (int l = 0; l < 3; l++)
          r[3 * s + l] = (double)x[3 * (s0 + s) + l] - (double)x[3 * t + l];
      mlp_batch<double>(w, r.data(), ns, y.data());
      for (int64_t s = 0; s < ns; s++)
        for (int row = 0; row < 3; row++)
          for (int col = 0; col < 3; col++)
            acc[row] += 0.5 * (y[9 * s + 3 * row + col] + y[9 * s + 3 * col + row]) *
                        (double)f[3 * (s0 + s) + col];
    }
    for (int d = 0; d < 3; d++) u_out[3 * q + d] = acc[d];
  });
}

// Factors of far block b (for the cross-interpolation tests): returns the
// rank k; C is k x |I| x 9, Q is k x |J| x 9, ipiv k entries, jpiv k+1.
int orc_model_aca_factors(void *h, const float *w, int64_t b, double *C, double *Q,
                          int64_t *ipiv, int64_t *jpiv) {
  Model *m = (Model *)h;
  AcaOut o;
  const int k = aca_block(*m, w, m->lists.farI[b], m->lists.farJ[b], nullptr, nullptr, &o);
  const int64_t mI = m->tr.size(m->lists.farI[b]), nJ = m->tr.size(m->lists.farJ[b]);
  for (int l = 0; l < k; l++) {
    if (C) std::memcpy(C + (int64_t)l * mI * 9, o.C[l].data(), sizeof(double) * mI * 9);
    if (Q) std::memcpy(Q + (int64_t)l * nJ * 9, o.Q[l].data(), sizeof(double) * nJ * 9);
    if (ipiv) ipiv[l] = o.ipiv[l];
  }
  if (jpiv)
    for (int l = 0; l <= k; l++) jpiv[l] = o.jpiv[l];
  return k;
}

// Stop-test quantities of far block b per iteration (k entries each): nu2
// (src/FarDot.cpp:948-1037) and mu2 after its update (:1037, :1041-1078).
int orc_model_aca_norms(void *h, const float *w, int64_t b, double *nu2, double *mu2) {
  Model *m = (Model *)h;
  AcaOut o;
  const int k = aca_block(*m, w, m->lists.farI[b], m->lists.farJ[b], nullptr, nullptr, &o);
  for (int l = 0; l < k; l++) {
    if (nu2) nu2[l] = o.nu2[l];
    if (mu2) mu2[l] = o.mu2[l];
  }
  return k;
}

// Bounded CPU-baseline sample (timing only): CloseDot over every
// close_stride-th close block and the full ACA + apply over every
// far_stride-th far block.  Returns the elapsed seconds; work[0] = near
// pair evaluations done, work[1] = far pair evaluations done
// (sum over sampled blocks of rank * (|I| + |J|)).
double orc_model_sample_time(void *h, const float *w, const float *f, int64_t close_stride,
                             int64_t far_stride, int64_t *work) {
  Model *m = (Model *)h;
  if (!m->built) orc_model_update(h, nullptr);
  std::vector<double> fp(3 * m->n), up(3 * m->n, 0.0);
  for (int64_t i = 0; i < m->n; i++)
    for (int d = 0; d < 3; d++) fp[3 * i + d] = f[3 * m->tr.map[i] + d];
  Model sub = *m;
  sub.lists.closeI.clear();
  sub.lists.closeJ.clear();
  sub.lists.farI.clear();
  sub.lists.farJ.clear();
  int64_t wn = 0;
  for (size_t b = 0; b < m->lists.closeI.size(); b += (size_t)close_stride) {
    sub.lists.closeI.push_back(m->lists.closeI[b]);
    sub.lists.closeJ.push_back(m->lists.closeJ[b]);
    wn += m->tr.size(m->lists.closeI[b]) * m->tr.size(m->lists.closeJ[b]);
  }
  for (size_t b = 0; b < m->lists.farI.size(); b += (size_t)far_stride) {
    sub.lists.farI.push_back(m->lists.farI[b]);
    sub.lists.farJ.push_back(m->lists.farJ[b]);
  }
  const auto t0 = std::chrono::steady_clock::now();
  close_dot(sub, w, fp.data(), up.data());
  far_dot(sub, w, fp.data(), up.data());
  const auto t1 = std::chrono::steady_clock::now();
  int64_t wf = 0;
  for (size_t b = 0; b < sub.lists.farI.size(); b++)
    wf += (int64_t)sub.far_rank[b] * (sub.tr.size(sub.lists.farI[b]) + sub.tr.size(sub.lists.farJ[b]));
  if (work) {
    work[0] = wn;
    work[1] = wf;
  }
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
