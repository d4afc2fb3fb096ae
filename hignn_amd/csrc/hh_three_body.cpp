// Neighbour lists and three-body / self corrections (C ABI of
// include/hh_three_body.h).
//
// Lists are built on the device from a hashed cell list (hh_forcing.hip):
// neighbour counts, a scan, the sorted neighbour lists (CSR), then the
// triplets (i, n, p) by the same count / scan / fill.  Only the totals come
// back to the host, to size the edge arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/hh_three_body.h"
#include "hh_cells.h"
#include "hh_forcing.h"
#include "hh_host.h"
#include "hh_three_body.h"

using namespace hh;

namespace {

// Neighbour and triplet lists of n particles at device positions x.
struct Lists {
  CellBufs cells;
  DBuf cnt, off, idx;           // neighbour CSR: counts, offsets (n + 1), indices
  DBuf tcnt, toff, trip;        // triplet counts, offsets (n + 1), (i, n, p) triples
  DBuf self;                    // self edges (i, n)
  int64_t n = 0, nnb = 0, ntrip = 0;
  std::vector<int32_t> src_tgt_h;  // image -> target (periodic two-body export)

  ForcingLaunch launch(const float *x, int64_t nn, float eps, float margin, bool periodic,
                       const float *lo, const float *hi) {
    ForcingLaunch L{};
    L.x = x;
    L.n = nn;
    L.eps = eps;
    L.margin = margin;
    L.periodic = periodic ? 1 : 0;
    for (int j = 0; j < 3; j++) L.lo[j] = lo[j], L.hi[j] = hi[j];
    cells.setup(L, nn);
    return L;
  }

  // Neighbour CSR within eps.
  void neighbours(const ForcingLaunch &L, hipStream_t st) {
    n = L.n;
    cnt.reserve(4 * (size_t)std::max<int64_t>(n, 1));
    off.reserve(4 * (size_t)(n + 1));
    HIPCHK(launch_neighbour_count(L, cnt.as<uint32_t>(), off.as<uint32_t>(), cells.bsum.as<uint32_t>(),
                                  st));
    std::vector<uint32_t> c((size_t)n);
    HIPCHK(hipMemcpyAsync(c.data(), cnt.p, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    uint64_t total = 0;
    for (uint32_t v : c) total += v;
    if (total > 0x7fffffffull)  // offsets are 32-bit, self edges index with int32
      throw HHError(HH_ERR_ARG, "more than 2^31 neighbour pairs: use a smaller radius or fewer particles");
    nnb = (int64_t)total;
    idx.reserve(4 * (size_t)std::max<int64_t>(nnb, 1));
    HIPCHK(launch_neighbour_fill(L, off.as<uint32_t>(), idx.as<int32_t>(), st));
  }

  // Triplets and self edges from the neighbour CSR.
  void triplets(hipStream_t st) {
    tcnt.reserve(4 * (size_t)std::max<int64_t>(n, 1));
    toff.reserve(4 * (size_t)(n + 1));
    HIPCHK(launch_triplet_count(off.as<uint32_t>(), idx.as<int32_t>(), n, tcnt.as<uint32_t>(), st));
    {  // offsets are 32-bit: refuse lists that would wrap
      std::vector<uint32_t> c((size_t)n);
      HIPCHK(hipMemcpyAsync(c.data(), tcnt.p, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      uint64_t tot = 0;
      for (uint32_t v : c) tot += v;
      if (tot > 0xffffffffull)
        throw HHError(HH_ERR_ARG, "more than 2^32 three-body edges: use a smaller radius or fewer particles");
    }
    HIPCHK(exclusive_scan(tcnt.as<uint32_t>(), n, toff.as<uint32_t>(), cells.bsum.as<uint32_t>(), st));
    uint32_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, toff.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    ntrip = total;
    trip.reserve(12 * (size_t)std::max<int64_t>(ntrip, 1));
    HIPCHK(launch_triplet_fill(off.as<uint32_t>(), idx.as<int32_t>(), n, toff.as<uint32_t>(),
                               trip.as<int32_t>(), st));
    // self edges (i, n) in CSR order
    std::vector<uint32_t> oh((size_t)n + 1);
    std::vector<int32_t> ih((size_t)nnb), eh(2 * (size_t)nnb);
    HIPCHK(hipMemcpyAsync(oh.data(), off.p, 4 * (size_t)(n + 1), hipMemcpyDeviceToHost, st));
    if (nnb) HIPCHK(hipMemcpyAsync(ih.data(), idx.p, 4 * (size_t)nnb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n; i++)
      for (uint32_t a = oh[i]; a < oh[i + 1]; a++) eh[2 * a] = (int32_t)i, eh[2 * a + 1] = ih[a];
    self.reserve(8 * (size_t)std::max<int64_t>(nnb, 1));
    if (nnb) HIPCHK(hipMemcpyAsync(self.p, eh.data(), 8 * (size_t)nnb, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }

  void release() {
    cells.release();
    for (DBuf *d : {&cnt, &off, &idx, &tcnt, &toff, &trip, &self}) d->release();
  }
};

constexpr int64_t kBatch = 1 << 18;  // edges per MLP batch
constexpr double kActScale = 2.8853900817779268;  // 2 log2(e)

size_t net_floats(int din) {
  return 256 * (size_t)din + 256 + 1024 * 256 + 1024 + 512 * 1024 + 512 + 9 * 512 + 9;
}

struct Net {
  DBuf w1s, b1s, w2s, b2s, w3s, b3s, w2h, w3h, w4, b4;
  int din = 0;

  void load(const float *w, int d) {
    din = d;
    const float *p = w;
    auto take = [&](size_t cnt) {
      const float *q = p;
      p += cnt;
      return q;
    };
    const float *W1 = take(256 * (size_t)d), *B1 = take(256), *W2 = take(1024 * 256),
                *B2 = take(1024), *W3 = take(512 * 1024), *B3 = take(512), *W4 = take(9 * 512),
                *B4 = take(9);
    auto up_scaled = [](DBuf &b, const float *src, size_t cnt, double s) {
      std::vector<float> v(cnt);
      for (size_t k = 0; k < cnt; k++) v[k] = (float)(s * (double)src[k]);
      b.reserve(4 * cnt);
      HIPCHK(hipMemcpy(b.p, v.data(), 4 * cnt, hipMemcpyHostToDevice));
    };
    auto up_half = [](DBuf &b, const float *src, size_t cnt, double s) {
      std::vector<_Float16> v(cnt);
      for (size_t k = 0; k < cnt; k++) v[k] = (_Float16)(float)(s * (double)src[k]);
      b.reserve(2 * cnt);
      HIPCHK(hipMemcpy(b.p, v.data(), 2 * cnt, hipMemcpyHostToDevice));
    };
    up_scaled(w1s, W1, 256 * (size_t)d, kActScale);
    up_scaled(b1s, B1, 256, kActScale);
    up_scaled(w2s, W2, 1024 * 256, kActScale);
    up_scaled(b2s, B2, 1024, kActScale);
    up_scaled(w3s, W3, 512 * 1024, kActScale);
    up_scaled(b3s, B3, 512, kActScale);
    up_half(w2h, W2, 1024 * 256, kActScale);
    up_half(w3h, W3, 512 * 1024, kActScale);
    up_scaled(w4, W4, 9 * 512, 1.0);
    up_scaled(b4, B4, 9, 1.0);
  }

  NetDev dev() const {
    return NetDev{w1s.as<float>(), b1s.as<float>(), w2s.as<float>(), b2s.as<float>(),
                  w3s.as<float>(), b3s.as<float>(), w2h.p,           w3h.p,
                  w4.as<float>(),  b4.as<float>()};
  }

  void release() {
    for (DBuf *b : {&w1s, &b1s, &w2s, &b2s, &w3s, &b3s, &w2h, &w3h, &w4, &b4}) b->release();
  }
};

}  // namespace

struct hh_neighbors {
  float eps2b = 0.0f, eps3b = 0.0f;
  bool periodic = false, has_domain = false;
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  DBuf x;
  int64_t n = 0;
  Lists two, three;
  bool built2 = false, built3 = false;

  void check_domain() const {
    if (periodic && !has_domain) throw HHError(HH_ERR_STATE, "periodic lists without a domain");
  }
  ~hh_neighbors() {
    two.release();
    three.release();
    x.release();
  }
};

struct hh_three_body {
  Net net3, nets;
  bool loaded = false;
  float eps = 0.0f;
  int prec = HH_PREC_FP32;
  Lists lists;
  DBuf u64, h1, h2, h3, rec, frc, xh, fh, uh;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int64_t last_edges[2] = {0, 0};
  double last_ms[2] = {0, 0};

  ~hh_three_body() {
    net3.release();
    nets.release();
    lists.release();
    for (DBuf *b : {&u64, &h1, &h2, &h3, &rec, &frc, &xh, &fh, &uh}) b->release();
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
  }

  void run_edges(const Net &net, const int32_t *edges, int64_t count, const float *x,
                 const float *f, hipStream_t st) {
    for (int64_t first = 0; first < count; first += kBatch) {
      EdgeBatch B{};
      B.x = x;
      B.f = f;
      B.edges = edges;
      B.first = first;
      B.count = std::min<int64_t>(kBatch, count - first);
      B.rec = rec.as<float4>();
      B.frc = frc.as<float4>();
      HIPCHK(launch_edge_mlp(B, net.dev(), net.din, prec == HH_PREC_FP16 ? 1 : 0, h1.p, h2.p, h3.p,
                             u64.as<double>(), st));
    }
  }

  void velocity(const float *x, const float *f, int64_t n, float *u, bool accumulate,
                hipStream_t st) {
    if (!loaded) throw HHError(HH_ERR_STATE, "three-body weights not loaded");
    if (!(eps > 0)) throw HHError(HH_ERR_STATE, "three-body radius not set");
    const float zero[3] = {0, 0, 0};
    ForcingLaunch L = lists.launch(x, n, eps, 0.0f, false, zero, zero);
    lists.neighbours(L, st);
    lists.triplets(st);
    const size_t es = prec == HH_PREC_FP16 ? 2 : 4;
    const int64_t bmax = std::min<int64_t>(kBatch, std::max(lists.ntrip, lists.nnb));
    const int64_t bpad = (std::max<int64_t>(bmax, 1) + 255) / 256 * 256;  // GEMM tile rows
    h1.reserve(es * 256 * (size_t)bpad);
    h2.reserve(es * 1024 * (size_t)bpad);
    h3.reserve(es * 512 * (size_t)bpad);
    rec.reserve(16 * (size_t)bpad);
    frc.reserve(16 * (size_t)bpad);
    u64.reserve(24 * (size_t)n);
    for (auto &e : ev)
      if (!e) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipMemsetAsync(u64.p, 0, 24 * (size_t)n, st));
    HIPCHK(hipEventRecord(ev[0], st));
    run_edges(net3, lists.trip.as<int32_t>(), lists.ntrip, x, f, st);
    HIPCHK(hipEventRecord(ev[1], st));
    HIPCHK(hipEventRecord(ev[2], st));
    run_edges(nets, lists.self.as<int32_t>(), lists.nnb, x, f, st);
    HIPCHK(hipEventRecord(ev[3], st));
    HIPCHK(launch_to_float(u64.as<double>(), u, 3 * n, accumulate ? 1 : 0, st));
    HIPCHK(hipEventSynchronize(ev[3]));
    float a = 0, b = 0;
    HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
    HIPCHK(hipEventElapsedTime(&b, ev[2], ev[3]));
    last_ms[0] = a;
    last_ms[1] = b;
    last_edges[0] = lists.ntrip;
    last_edges[1] = lists.nnb;
  }
};

namespace {

template <class T>
T *need(T *p, const char *what) {
  if (!p) throw HHError(HH_ERR_ARG, std::string("NULL ") + what);
  return p;
}

// Builds the lists of nb within eps (images with the two-body margin).
void build_lists(hh_neighbors *nb, Lists &l, float eps, bool three) {
  if (nb->n <= 0) throw HHError(HH_ERR_STATE, "update_coord has not run");
  nb->check_domain();
  ForcingLaunch L = l.launch(nb->x.as<float>(), nb->n, eps, nb->eps2b, nb->periodic, nb->lo, nb->hi);
  l.neighbours(L, nullptr);
  if (three) l.triplets(nullptr);
  // image -> target map for the two-body export (a second lookup through it
  // is what BuildTwoBodyInfo stores, NeighborLists.hpp:497-500)
  uint32_t ns = (uint32_t)nb->n;
  if (nb->periodic)
    HIPCHK(hipMemcpy(&ns, l.cells.off.as<uint32_t>() + nb->n, 4, hipMemcpyDeviceToHost));
  l.src_tgt_h.resize(ns);
  HIPCHK(hipMemcpy(l.src_tgt_h.data(), l.cells.src_tgt.p, 4 * (size_t)ns, hipMemcpyDeviceToHost));
  HIPCHK(hipDeviceSynchronize());
}

}  // namespace

extern "C" {

int hh_neighbors_create(hh_neighbors **out) {
  return guarded([&]() {
    need(out, "argument");
    ensure_device();
    *out = new hh_neighbors();
  });
}

void hh_neighbors_destroy(hh_neighbors *nb) { delete nb; }

int hh_neighbors_set_two_body_epsilon(hh_neighbors *nb, float epsilon) {
  return guarded([&]() { need(nb, "lists")->eps2b = epsilon; });
}

int hh_neighbors_set_three_body_epsilon(hh_neighbors *nb, float epsilon) {
  return guarded([&]() { need(nb, "lists")->eps3b = epsilon; });
}

int hh_neighbors_set_periodic(hh_neighbors *nb, int flag) {
  return guarded([&]() { need(nb, "lists")->periodic = flag != 0; });
}

int hh_neighbors_set_domain(hh_neighbors *nb, const float lo[3], const float hi[3]) {
  return guarded([&]() {
    need(nb, "lists");
    need(lo, "domain");
    need(hi, "domain");
    for (int j = 0; j < 3; j++) nb->lo[j] = lo[j], nb->hi[j] = hi[j];
    nb->periodic = nb->has_domain = true;
  });
}

int hh_neighbors_update_coord(hh_neighbors *nb, const float *x, int64_t n) {
  return guarded([&]() {
    need(nb, "lists");
    need(x, "coordinates");
    if (n <= 0) throw HHError(HH_ERR_ARG, "no particles");
    nb->x.reserve(12 * (size_t)n);
    HIPCHK(hipMemcpy(nb->x.p, x, 12 * (size_t)n, hipMemcpyHostToDevice));
    nb->n = n;
    nb->built2 = nb->built3 = false;
  });
}

int hh_neighbors_build_two_body(hh_neighbors *nb, int64_t *num_edges) {
  return guarded([&]() {
    need(nb, "lists");
    if (!(nb->eps2b > 0)) throw HHError(HH_ERR_STATE, "two-body radius not set");
    build_lists(nb, nb->two, nb->eps2b, false);
    nb->built2 = true;
    if (num_edges) *num_edges = nb->two.nnb;
  });
}

int hh_neighbors_build_three_body(hh_neighbors *nb, int64_t *num_edges, int64_t *num_self_edges) {
  return guarded([&]() {
    need(nb, "lists");
    if (!(nb->eps3b > 0)) throw HHError(HH_ERR_STATE, "three-body radius not set");
    build_lists(nb, nb->three, nb->eps3b, true);
    nb->built3 = true;
    if (num_edges) *num_edges = nb->three.ntrip;
    if (num_self_edges) *num_self_edges = nb->three.nnb;
  });
}

int hh_neighbors_two_body_edges(hh_neighbors *nb, int64_t *out) {
  return guarded([&]() {
    need(nb, "lists");
    need(out, "output");
    if (!nb->built2) throw HHError(HH_ERR_STATE, "build_two_body_info has not run");
    const Lists &l = nb->two;
    std::vector<uint32_t> oh((size_t)l.n + 1);
    std::vector<int32_t> ih((size_t)l.nnb);
    HIPCHK(hipMemcpy(oh.data(), l.off.p, 4 * ((size_t)l.n + 1), hipMemcpyDeviceToHost));
    if (l.nnb) HIPCHK(hipMemcpy(ih.data(), l.idx.p, 4 * (size_t)l.nnb, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < l.n; i++)
      for (uint32_t a = oh[i]; a < oh[i + 1]; a++) {
        out[a] = i;
        out[l.nnb + a] = l.src_tgt_h[(size_t)ih[a]];
      }
  });
}

int hh_neighbors_three_body_edges(hh_neighbors *nb, int64_t *out) {
  return guarded([&]() {
    need(nb, "lists");
    need(out, "output");
    if (!nb->built3) throw HHError(HH_ERR_STATE, "build_three_body_info has not run");
    const Lists &l = nb->three;
    std::vector<int32_t> t(3 * (size_t)l.ntrip);
    if (l.ntrip) HIPCHK(hipMemcpy(t.data(), l.trip.p, 12 * (size_t)l.ntrip, hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < l.ntrip; e++)
      for (int r = 0; r < 3; r++) out[r * l.ntrip + e] = t[3 * (size_t)e + r];
  });
}

int hh_neighbors_three_body_self_edges(hh_neighbors *nb, int64_t *out) {
  return guarded([&]() {
    need(nb, "lists");
    need(out, "output");
    if (!nb->built3) throw HHError(HH_ERR_STATE, "build_three_body_info has not run");
    const Lists &l = nb->three;
    std::vector<int32_t> t(2 * (size_t)l.nnb);
    if (l.nnb) HIPCHK(hipMemcpy(t.data(), l.self.p, 8 * (size_t)l.nnb, hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < l.nnb; e++)
      for (int r = 0; r < 2; r++) out[r * l.nnb + e] = t[2 * (size_t)e + r];
  });
}

int hh_three_body_create(hh_three_body **out) {
  return guarded([&]() {
    need(out, "argument");
    ensure_device();
    *out = new hh_three_body();
  });
}

void hh_three_body_destroy(hh_three_body *tb) { delete tb; }

int hh_three_body_load(hh_three_body *tb, const float *w3, int64_t n3, const float *wself,
                       int64_t nself) {
  return guarded([&]() {
    need(tb, "model");
    need(w3, "weights");
    need(wself, "weights");
    if ((size_t)n3 != net_floats(6) || (size_t)nself != net_floats(3))
      throw HHError(HH_ERR_ARG, "expected " + std::to_string(net_floats(6)) + " three-body and " +
                                    std::to_string(net_floats(3)) + " self-net floats");
    tb->net3.load(w3, 6);
    tb->nets.load(wself, 3);
    tb->loaded = true;
  });
}

int hh_three_body_set_epsilon(hh_three_body *tb, float epsilon) {
  return guarded([&]() { need(tb, "model")->eps = epsilon; });
}

int hh_three_body_set_precision(hh_three_body *tb, int prec) {
  return guarded([&]() {
    if (prec != HH_PREC_FP32 && prec != HH_PREC_FP16)
      throw HHError(HH_ERR_ARG, "three-body precision must be FP32 or FP16");
    need(tb, "model")->prec = prec;
  });
}

int hh_three_body_velocity_device(hh_three_body *tb, const float *x_dev, const float *f_dev,
                                  int64_t n, float *u_dev, int accumulate, void *stream) {
  return guarded([&]() {
    need(tb, "model");
    if (!x_dev || !f_dev || !u_dev || n <= 0) throw HHError(HH_ERR_ARG, "invalid arrays");
    tb->velocity(x_dev, f_dev, n, u_dev, accumulate != 0, (hipStream_t)stream);
  });
}

int hh_three_body_velocity(hh_three_body *tb, const float *x, const float *f, int64_t n,
                           float *u) {
  return guarded([&]() {
    need(tb, "model");
    if (!x || !f || !u || n <= 0) throw HHError(HH_ERR_ARG, "invalid arrays");
    const size_t b = 12 * (size_t)n;
    tb->xh.reserve(b);
    tb->fh.reserve(b);
    tb->uh.reserve(b);
    HIPCHK(hipMemcpy(tb->xh.p, x, b, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(tb->fh.p, f, b, hipMemcpyHostToDevice));
    tb->velocity(tb->xh.as<float>(), tb->fh.as<float>(), n, tb->uh.as<float>(), false, nullptr);
    HIPCHK(hipMemcpy(u, tb->uh.p, b, hipMemcpyDeviceToHost));
  });
}

int hh_three_body_last_profile(hh_three_body *tb, int64_t edges[2], double ms[2]) {
  return guarded([&]() {
    need(tb, "model");
    if (edges) edges[0] = tb->last_edges[0], edges[1] = tb->last_edges[1];
    if (ms) ms[0] = tb->last_ms[0], ms[1] = tb->last_ms[1];
  });
}

}  // extern "C"
