// libhh: model object, orchestration and the C ABI declared in include/hh.h.
//
// Mirrors HignnModel (include/HignnModel.hpp, src/HignnModel.cpp): the model
// owns the coordinates, the tree and block lists built by update_coord, and
// device copies of everything dot needs.  dot runs
//   gather f into tree order -> near field (close leaf pairs) -> far field
//   (batched block ACA + low-rank apply) -> scatter u to the original order,
// all on one HIP stream.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include <thread>
#include <atomic>

#include "../../include/hh.h"
#include "../../include/hh_dynamics.h"
#include "../../include/hh_device.h"
#include "../../include/hh_surrogate.h"
#include "hh_host.h"
#include "hh_internal.h"
#include "hh_kernels.h"
#include "hh_surrogate_internal.h"

using namespace hh;

namespace hh {

thread_local std::string g_last_error;
int g_device = -1;
int g_cus = 256;

void ensure_device() {
  if (g_device >= 0) return;
  int count = 0;
  HIPCHK(hipGetDeviceCount(&count));
  if (count <= 0) throw HHError(HH_ERR_DEVICE, "no HIP device");
  int dev = 0;
  if (const char *lr = std::getenv("LOCAL_RANK")) dev = std::atoi(lr) % count;
  HIPCHK(hipSetDevice(dev));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  g_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  g_device = dev;
}

}  // namespace hh

namespace {

uint16_t to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

uint16_t to_f16(float f) {
  _Float16 h = (_Float16)f;
  uint16_t u;
  std::memcpy(&u, &h, 2);
  return u;
}

}  // namespace

struct hh_model {
  int64_t n = 0;
  int leaf = 0;
  std::vector<float> x;  // original order
  // parameters (src/HignnModel.cpp:335-386 defaults)
  double epsilon = 0.05, eta = 1.0, max_far_dist = 1000.0;
  int max_iter = 100, pool_factor = 40, far_work = 5000;
  int64_t max_rel_coord = 500000;
  bool post_check = false, use_symmetry = true;
  bool tree_reuse = false;  // update_coord keeps the last tree and lists (opt-in)
  // tree and block lists built on the device (hh_tree_gpu.hip, the default)
  // or on host threads (hh_tree.cpp); both give the same result bit for bit
  bool tree_device = true;
  DeviceTreeBuilder *tbuild = nullptr;
  DBuf xdev;  // coordinates (n,3) uploaded for a device build from host data
  uint64_t builds = 0, loads = 0;  // tree builds and weight loads (graph signatures)
  // independent instances (hh_model_create_instances): particle counts, in
  // row order of x; empty for a single system
  std::vector<int64_t> inst_counts;
  int prec = PREC_FP32;
  // byte budget of one far-field batch (factor pools + Gram histories); 0:
  // 70% of the device memory free at the first far field, leaving at least
  // 32 GiB (fewer, larger batches: each batch runs its own iteration loop)
  size_t far_pool_bytes = 0;
  hipStream_t stream = nullptr;
  // weights
  bool has_weights = false;
  std::vector<float> w;
  DBuf wf32[5], wf64[5], wh_f[3], wh_bf16[2], wh_f16[2];
  // tree + lists
  TreeLists tl;
  bool built = false;
  DBuf xp, perm, fp, u, fin, uout;
  DBuf near_blocks, near_tasks, near_next;
  int64_t n_near_tasks = 0;
  int64_t near_pairs = 0;  // pair-function evaluations of the near field
  DBuf dense_blocks, dense_tasks;
  int64_t n_dense_tasks = 0;
  std::vector<FarBlock> far;  // is, m, js, n (offsets set per batch)
  std::vector<int32_t> far_rank;
  // far workspace of the batched iterations: blocks with pool offsets (all
  // batches, far-list order), per-block ACA state, per-slot lists of the
  // current iteration (active list double-buffered for the device plan),
  // pass tile lists per kind, chunk partials, apply lists, counts
  DBuf fb_all, it, ipiv, jpiv, mu2, nu2, stop, act2[2], ybase_c, ybase_r, applylist, Y, Cpool, Qpool;
  DBuf toff_c, toff_r, tiles_fc, tiles_sc, tiles_fr, tiles_sr, cfirst_c, cfirst_r, part_c, part_r,
      x_c, x_r, invb;
  DBuf a_first_i, a_first_j, w_i, w_j, cnt2[2], plan_part, rank_dev, ent_acc;
  PBuf h_stop, h_cnt, h_cnt_apply, h_act, h_ybase_c, h_ybase_r, h_toff_c, h_toff_r, h_cfirst_c,
      h_cfirst_r, h_list, h_afirst_i, h_afirst_j;
  // batches of the batched far field (blocks whose factor pools fit the byte
  // budget) and the bounds of their per-iteration counts (every block active)
  struct FarBatch {
    int64_t b0, b1;
    int64_t ntc, ntr, ncc, ncr, ysum_c, ysum_r;
  };
  std::vector<FarBatch> batches;
  uint64_t batch_key = 0;  // configuration the batches were laid out for
  // 1 (default): each iteration's active list, prefix counts and stop flags
  // stay on the device (aca_plan kernels), no host round trip per iteration
  // (HH_ACA_CONTROL=host: the host compacts after reading the stop flags)
  int aca_control = 1;
  bool judge_pending = false;  // judge flags / errors / entry counts not yet read back
  std::vector<hipEvent_t> evpool;  // events around the full-MLP ACA passes
  // persistent far-field workspace
  // far field: 1 batched iterations, 0 persistent kernel, -1 (default) the
  // persistent kernel when there are no more far blocks than compute units
  // (one workgroup per block, no per-iteration launches), batched otherwise
  int far_mode = -1;
  // 16-bit modes: ACA factor pools stored in BF16 (arithmetic in FP32,
  // norms and sums in FP64); HH_FAR_C16=0 / HH_FAR_Q16=0 keep the C / Q pool
  // in FP32
  bool far_c16 = true, far_q16 = true;
  // BF16 pools only for far fields on full-MLP entries: the 16-bit MLP's own
  // entry noise (about 1e-3) already keeps those ACAs from the tolerance
  // test, so their ranks run to max_iter either way.  Surrogate entries are
  // smooth FP32, the ACA converges, and factors rounded to BF16 (8-bit
  // mantissa) would leave a residual floor above epsilon: such blocks then
  // run to max_iter on rounding noise, with errors up to 1e-1 at single
  // targets measured at N = 1.5e5.  With a surrogate the pools are FP32.
  bool pool16(bool c) const { return prec != PREC_FP32 && !sur && (c ? far_c16 : far_q16); }
  bool verbose = false;  // HH_VERBOSE=1: far-field progress lines on stderr
  int aca_chunk = kAcaChunk;  // HH_ACA_CHUNK: items per ACA step chunk (1..kAcaChunk)
  int near_hcfg = 0;          // tile shape of the 16-bit near kernel (kNearShapes)
  int near_cus = 0;           // workgroups of the persistent 16-bit near kernel (0: one per CU)
  int mlp_cus = 0;            // workgroups of the 16-bit full-MLP ACA kernel (0: one per CU)
  DBuf p_blk, p_order, p_rank, p_poolC, p_poolQ, p_Y;
  bool p_ready = false;
  int64_t p_slotC = 0, p_slotQ = 0, p_slotY = 0;
  int p_nslots = 0, p_maxiter = 0, p_f64 = -1;
  // check workspace
  DBuf chk_blocks, chk_tasks, chk_tlist, chk_u;
  double t_close = 0, t_far = 0, t_total = 0;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  // pair-function kernels of the last dot: [0] near field, [1] ACA passes
  double prof_ms[2] = {0, 0};
  int64_t prof_pairs[2] = {0, 0}, prof_launches[2] = {0, 0};
  hipEvent_t evm[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t ev_all[2] = {nullptr, nullptr};  // whole matvec
  // batched far field: the near field runs beside it on the side stream
  // (HH_NEAR_OVERLAP=0 runs them one after the other)
  bool near_overlap = true;
  // persistent far field: the near field runs beside it on a second stream
  // (forked and joined with events), on the compute units the far blocks
  // leave idle
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr, far0_ev = nullptr;
  // far-field surrogate (borrowed; attached by hh_model_set_surrogate)
  hh_surrogate *sur = nullptr;
  double sur_tol = 2e-3;
  int sur_samples = 32;
  DBuf jd_blk, jd_r, jd_m9, jd_s6, jd_flag, jd_err;
  std::vector<int32_t> sur_flag;  // per far block: 1 = surrogate entries
  std::vector<float> sur_err;     // per far block: sampled relative error
  // [judged, on surrogate, fell back, surrogate entries, full-MLP far entries]
  int64_t sur_stats[5] = {0, 0, 0, 0, 0};

  ~hh_model() {
    if (tbuild) device_tree_builder_free(tbuild);
    xdev.release();
    for (auto *b : {&p_blk, &p_order, &p_rank, &p_poolC, &p_poolQ, &p_Y, &jd_blk, &jd_r, &jd_m9,
                    &jd_s6, &jd_flag, &jd_err})
      b->release();
    for (auto *b : {&xp, &perm, &fp, &u, &fin, &uout, &near_blocks, &near_tasks, &dense_blocks,
                    &dense_tasks, &fb_all, &it, &ipiv, &jpiv, &mu2, &nu2, &stop, &act2[0],
                    &act2[1], &ybase_c, &ybase_r, &applylist, &Y, &Cpool,
                    &Qpool, &chk_blocks, &chk_tasks, &chk_tlist, &chk_u, &toff_c, &toff_r, &tiles_fc,
                    &tiles_sc, &tiles_fr, &tiles_sr, &cfirst_c, &cfirst_r, &part_c, &part_r, &x_c, &x_r,
                    &invb, &a_first_i, &a_first_j, &w_i, &w_j, &cnt2[0], &cnt2[1], &plan_part,
                    &rank_dev, &ent_acc, &near_next})
      b->release();
    for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_na)
      if (e) (void)hipEventDestroy(e);
    h_na.release();
    for (auto &b : wf32) b.release();
    for (auto &b : wf64) b.release();
    for (auto &b : wh_f) b.release();
    for (auto &b : wh_bf16) b.release();
    for (auto &b : wh_f16) b.release();
    for (auto *b : {&h_stop, &h_cnt, &h_cnt_apply, &h_act, &h_ybase_c, &h_ybase_r, &h_toff_c,
                    &h_toff_r, &h_cfirst_c, &h_cfirst_r, &h_list, &h_afirst_i, &h_afirst_j})
      b->release();
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto &e : evm)
      if (e) (void)hipEventDestroy(e);
    for (auto &e : ev_all)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {fork_ev, join_ev, far0_ev})
      if (e) (void)hipEventDestroy(e);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void upload(DBuf &b, const void *src, size_t bytes) {
    b.reserve(bytes);
    if (bytes) HIPCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
  }

  // ---- weights in fragment order ----------------------------------------
  // Block entry e = 3 row + col carried by layer-3 output row d, or -1.  The
  // kernels read row `row` of a pair's block from lane group g = row,
  // register r = col, i.e. output row 4g + r (4r + g for the f64 MFMA).
  static int entry_of_row(int d, bool f64) {
    const int g = f64 ? d % 4 : d / 4, r = f64 ? d / 4 : d % 4;
    return (g < 3 && r < 3) ? 3 * g + r : -1;
  }
  // The FP32 and 16-bit kernels evaluate tanh x as 1 - 2/(2^x' + 1) with
  // x' = 2 x log2 e; scaling W1, b1, W2, b2 by 2 log2 e here saves the
  // multiply per activation.  The FP64 kernels use unscaled weights.
  static double act_scale(bool f64) { return f64 ? 1.0 : 2.0 * 1.4426950408889634; }

  template <class R>
  void prep_f(DBuf (&dst)[5]) {
    const float *W1 = w.data(), *B1 = W1 + 384, *W2 = B1 + 128, *B2 = W2 + 65536, *W3 = B2 + 512,
                *B3 = W3 + 4608;
    const bool f64 = sizeof(R) == 8;
    const double sc = act_scale(f64);
    std::vector<R> w1b(512), w2f(32 * 32 * 64), b2(512), w3f(32 * 4 * 64), b3(16, (R)0);
    for (int k = 0; k < 128; k++) {
      for (int d = 0; d < 3; d++) w1b[4 * k + d] = (R)(sc * W1[3 * k + d]);
      w1b[4 * k + 3] = (R)(sc * B1[k]);
    }
    for (int mt = 0; mt < 32; mt++)
      for (int s = 0; s < 32; s++)
        for (int l = 0; l < 64; l++) {
          const int p = l & 15, g = l >> 4;
          w2f[(mt * 32 + s) * 64 + l] = (R)(sc * W2[(16 * mt + p) * 128 + 4 * s + g]);
        }
    for (int k = 0; k < 512; k++) b2[k] = (R)(sc * B2[k]);
    // layer-2 output register r of lane group g is feature 16mt + row(g, r):
    // 4g + r for the f32 MFMA, 4r + g for the f64 MFMA
    for (int mt = 0; mt < 32; mt++)
      for (int r = 0; r < 4; r++)
        for (int l = 0; l < 64; l++) {
          const int p = l & 15, g = l >> 4;
          const int row = f64 ? 4 * r + g : 4 * g + r;
          const int e = entry_of_row(p, f64);
          w3f[(mt * 4 + r) * 64 + l] = e >= 0 ? (R)W3[e * 512 + 16 * mt + row] : (R)0;
        }
    for (int d = 0; d < 16; d++) {
      const int e = entry_of_row(d, f64);
      b3[d] = e >= 0 ? (R)B3[e] : (R)0;
    }
    upload(dst[0], w1b.data(), w1b.size() * sizeof(R));
    upload(dst[1], w2f.data(), w2f.size() * sizeof(R));
    upload(dst[2], b2.data(), b2.size() * sizeof(R));
    upload(dst[3], w3f.data(), w3f.size() * sizeof(R));
    upload(dst[4], b3.data(), b3.size() * sizeof(R));
  }

  void prep_h(DBuf (&dst)[2], uint16_t (*cvt)(float)) {
    const float *W2 = w.data() + 512, *W3 = W2 + 65536 + 512;
    const double sc = act_scale(false);
    std::vector<uint16_t> w2f(32 * 4 * 64 * 8), w3f(16 * 64 * 8);
    for (int mt = 0; mt < 32; mt++)
      for (int s = 0; s < 4; s++)
        for (int l = 0; l < 64; l++)
          for (int j = 0; j < 8; j++) {
            const int p = l & 15, g = l >> 4;
            w2f[((mt * 4 + s) * 64 + l) * 8 + j] =
                cvt((float)(sc * W2[(16 * mt + p) * 128 + 32 * s + 8 * g + j]));
          }
    for (int q = 0; q < 16; q++)
      for (int l = 0; l < 64; l++)
        for (int j = 0; j < 8; j++) {
          const int p = l & 15, g = l >> 4;
          const int feat = j < 4 ? 32 * q + 4 * g + j : 32 * q + 16 + 4 * g + (j - 4);
          const int e = entry_of_row(p, false);
          w3f[(q * 64 + l) * 8 + j] = e >= 0 ? cvt(W3[e * 512 + feat]) : (uint16_t)0;
        }
    upload(dst[0], w2f.data(), w2f.size() * 2);
    upload(dst[1], w3f.data(), w3f.size() * 2);
  }

  void load_weights(const float *wts) {
    w.assign(wts, wts + 71177);
    prep_f<float>(wf32);
    prep_f<double>(wf64);
    // 16-bit path: FP32 layer 1 (MFMA A fragments) and biases.  Fragment
    // (s, half), lane 16 g + p holds column g of [W1 | b1] (x, y, z, bias)
    // for hidden unit 32 s + 8 (p / 4) + 4 half + p % 4, so that the layer-1
    // output rows land in layer 2's operand order.
    const float *W1 = w.data(), *B1 = W1 + 384, *B2 = B1 + 128 + 65536, *B3 = B2 + 512 + 4608;
    const double sc = act_scale(false);
    std::vector<float> w1b(512), b2(512), b3(16, 0.0f);
    for (int s = 0; s < 4; s++)
      for (int half = 0; half < 2; half++)
        for (int lane = 0; lane < 64; lane++) {
          const int g = lane / 16, p = lane % 16;
          const int k = 32 * s + 8 * (p / 4) + 4 * half + p % 4;
          w1b[(size_t)(2 * s + half) * 64 + lane] = (float)(sc * (g < 3 ? W1[3 * k + g] : B1[k]));
        }
    for (int k = 0; k < 512; k++) b2[k] = (float)(sc * B2[k]);
    for (int d = 0; d < 16; d++) {
      const int e = entry_of_row(d, false);
      b3[d] = e >= 0 ? B3[e] : 0.0f;
    }
    upload(wh_f[0], w1b.data(), 512 * 4);
    upload(wh_f[1], b2.data(), 512 * 4);
    upload(wh_f[2], b3.data(), 16 * 4);
    prep_h(wh_bf16, to_bf16);
    prep_h(wh_f16, to_f16);
    has_weights = true;
    loads++;
  }

  DevWeightsF devw_f(bool f64) const {
    const DBuf *b = f64 ? wf64 : wf32;
    return DevWeightsF{b[0].p, b[1].p, b[2].p, b[3].p, b[4].p};
  }
  DevWeightsH devw_h(int p) const {
    const DBuf *b = p == PREC_FP16 ? wh_f16 : wh_bf16;
    return DevWeightsH{wh_f[0].as<float>(), wh_f[1].as<float>(), wh_f[2].as<float>(), b[0].p,
                       b[1].p};
  }

  // ---- update_coord ------------------------------------------------------
  // The host coordinates after device-side updates on a kept tree: the
  // positions live only in xp (tree order) until a rebuild needs them.
  bool x_stale = false;
  void refresh_host_x() {
    if (!x_stale) return;
    std::vector<float> xph(4 * (size_t)n);
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(xph.data(), xp.p, 16 * (size_t)n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
      const int64_t s = tl.perm[(size_t)i];
      for (int d = 0; d < 3; d++) x[3 * (size_t)s + d] = xph[4 * (size_t)i + d];
    }
    x_stale = false;
  }

  // xdev_in: the new coordinates (n,3, original order) in device memory, or
  // NULL when they are in x
  void update(const float *xdev_in = nullptr) {
    if (!(tree_reuse && built) && !xdev_in) refresh_host_x();
    if (tree_reuse && built) {
      // keep the permutation, tree and block lists of the last build; only
      // the positions (in that tree order) are refreshed
      std::vector<float> xph(4 * (size_t)n);
      for (int64_t i = 0; i < n; i++) {
        const int64_t s = tl.perm[(size_t)i];
        for (int d = 0; d < 3; d++) xph[4 * (size_t)i + d] = x[3 * (size_t)s + d];
        xph[4 * (size_t)i + 3] = 0.0f;
      }
      upload(xp, xph.data(), xph.size() * 4);
      return;
    }
    if (tree_device) {
      const float *xd = xdev_in;
      if (!xd) {
        xdev.reserve(12 * (size_t)n);
        HIPCHK(hipMemcpy(xdev.p, x.data(), 12 * (size_t)n, hipMemcpyHostToDevice));
        xd = xdev.as<float>();
      }
      std::vector<int64_t> pb{0};
      if (inst_counts.size() > 1)
        for (int64_t c : inst_counts) pb.push_back(pb.back() + c);
      else
        pb.push_back(n);
      if (!tbuild) tbuild = device_tree_builder_new();
      perm.reserve(4 * (size_t)n);
      xp.reserve(16 * (size_t)n);
      build_tree_lists_device(tbuild, xd, pb, leaf, use_symmetry, max_rel_coord, stream, tl,
                              perm.as<int32_t>());
      // positions in tree order
      HIPCHK(launch_gather_forces(xd, n, perm.as<int32_t>(), xp.as<float4>(), stream));
      HIPCHK(hipStreamSynchronize(stream));
      if (xdev_in) x_stale = true;  // the host copy catches up when a host path needs it
    } else {
      if (inst_counts.size() > 1)
        build_instances();
      else
        build_tree_lists(x.data(), n, leaf, use_symmetry, max_rel_coord, 0, tl);
      // positions in tree order
      std::vector<float> xph(4 * (size_t)n);
      std::vector<int32_t> ph((size_t)n);
      for (int64_t i = 0; i < n; i++) {
        const int64_t s = tl.perm[(size_t)i];
        ph[(size_t)i] = (int32_t)s;
        for (int d = 0; d < 3; d++) xph[4 * (size_t)i + d] = x[3 * (size_t)s + d];
        xph[4 * (size_t)i + 3] = 0.0f;
      }
      upload(xp, xph.data(), xph.size() * 4);
      upload(perm, ph.data(), ph.size() * 4);
    }
    builds++;
    // near-field row blocks and tasks (src/CloseDot.cpp)
    std::vector<RowBlock> rb;
    std::vector<NearTask> tk;
    near_pairs = 0;
    for (size_t b = 0; b < tl.closeI.size(); b++) {
      const int64_t I = tl.closeI[b], J = tl.closeJ[b];
      RowBlock r{(int32_t)tl.start(I), (int32_t)tl.size(I), (int32_t)tl.start(J),
                 (int32_t)tl.size(J), (use_symmetry && J > I) ? BF_REUSE : 0, 0, 0, 0};
      rb.push_back(r);
      near_pairs += (int64_t)r.m * r.n;
      for (int32_t o = 0; o < r.m; o += 4) tk.push_back(NearTask{(int32_t)b, o});
    }
    upload(near_blocks, rb.data(), rb.size() * sizeof(RowBlock));
    upload(near_tasks, tk.data(), tk.size() * sizeof(NearTask));
    n_near_tasks = (int64_t)tk.size();
    n_dense_tasks = -1;  // built on demand
    far.clear();
    for (size_t b = 0; b < tl.farI.size(); b++) {
      const int64_t I = tl.farI[b], J = tl.farJ[b];
      far.push_back(FarBlock{(int32_t)tl.start(I), (int32_t)tl.size(I), (int32_t)tl.start(J),
                             (int32_t)tl.size(J), 0, 0});
    }
    far_rank.assign(far.size(), 0);
    rank_pending = judge_pending = false;
    batches.clear();
    p_ready = false;
    built = true;
  }

  // Trees and block lists of independent instances, built per instance on
  // host threads and merged: instance k's node ids move up by the nodes of
  // instances 0..k-1 and its particle ranges by their particles, so the
  // merged lists drive one near-field launch and one far field over all
  // instances with no block between two instances.  Leaves keep child id 0.
  void build_instances() {
    const size_t K = inst_counts.size();
    std::vector<TreeLists> parts(K);
    std::vector<int64_t> pbase(K + 1, 0);
    for (size_t k = 0; k < K; k++) pbase[k + 1] = pbase[k] + inst_counts[k];
    const int nt = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<size_t> next(0);
    std::vector<std::thread> th;
    std::vector<std::string> errs(K);
    for (int t = 0; t < std::min<int>(nt, (int)K); t++)
      th.emplace_back([&]() {
        for (size_t k; (k = next.fetch_add(1)) < K;) {
          try {
            build_tree_lists(x.data() + 3 * pbase[k], inst_counts[k], leaf, use_symmetry,
                             max_rel_coord, 1, parts[k]);
          } catch (const std::exception &e) {
            errs[k] = e.what();
          }
        }
      });
    for (auto &t : th) t.join();
    for (size_t k = 0; k < K; k++)
      if (!errs[k].empty()) throw HHError(HH_ERR_ARG, "instance " + std::to_string(k) + ": " + errs[k]);
    TreeLists out;
    out.n = n;
    int64_t nbase = 0;
    for (size_t k = 0; k < K; k++) {
      const TreeLists &p = parts[k];
      const uint64_t nb = (uint64_t)nbase, pb = (uint64_t)pbase[k];
      for (int64_t nd = 0; nd < p.T; nd++) {
        const uint64_t *r = &p.rows[5 * (size_t)nd];
        out.rows.push_back(r[0] ? r[0] + nb : 0);
        out.rows.push_back(r[0] ? r[1] + nb : r[1]);
        out.rows.push_back(r[2] + pb);
        out.rows.push_back(r[3] + pb);
        out.rows.push_back(r[4]);
      }
      out.aux.insert(out.aux.end(), p.aux.begin(), p.aux.end());
      for (int64_t v : p.perm) out.perm.push_back(v + pbase[k]);
      for (int64_t v : p.leaves) out.leaves.push_back(v + nbase);
      for (int64_t v : p.closeI) out.closeI.push_back(v + nbase);
      for (int64_t v : p.closeJ) out.closeJ.push_back(v + nbase);
      for (int64_t v : p.farI) out.farI.push_back(v + nbase);
      for (int64_t v : p.farJ) out.farJ.push_back(v + nbase);
      nbase += p.T;
    }
    out.T = nbase;
    tl = std::move(out);
  }

  void ensure_buffers() {
    fp.reserve(16 * (size_t)n);
    u.reserve(24 * (size_t)n);
    fin.reserve(12 * (size_t)n);
    uout.reserve(12 * (size_t)n);
    near_next.reserve(256);
  }

  // ---- far field: persistent kernel -------------------------------------
  // Blocks go largest first (cost k_max (m + n)); each workgroup owns one
  // slot of the C/Q pools sized for the largest block.
  void far_field_persistent(hipStream_t st) {
    const int64_t nfar = (int64_t)far.size();
    if (nfar == 0) return;
    const bool f64 = prec == PREC_FP32;
    const size_t esz = (f64 ? 8 : 4) * 9;
    const int MI = max_iter;
    if (!p_ready || p_maxiter != MI || p_f64 != (int)f64) {
      std::vector<int32_t> order((size_t)nfar);
      for (int64_t b = 0; b < nfar; b++) order[(size_t)b] = (int32_t)b;
      auto cost = [&](int32_t b) {
        const FarBlock &f = far[(size_t)b];
        return (int64_t)std::min<int64_t>(MI, std::min(f.m, f.n)) * (f.m + f.n);
      };
      std::stable_sort(order.begin(), order.end(),
                       [&](int32_t a, int32_t b) { return cost(a) > cost(b); });
      std::vector<FarBlock> hb((size_t)nfar);
      int64_t sc = 1, sq = 1, sy = 1;
      for (int64_t p = 0; p < nfar; p++) {
        hb[(size_t)p] = far[(size_t)order[(size_t)p]];
        const FarBlock &f = hb[(size_t)p];
        const int64_t kmax = std::min<int64_t>(MI, std::min(f.m, f.n));
        sc = std::max(sc, kmax * f.m);
        sq = std::max(sq, kmax * f.n);
        sy = std::max<int64_t>(sy, std::max(f.m, f.n));
      }
      upload(p_blk, hb.data(), hb.size() * sizeof(FarBlock));
      upload(p_order, order.data(), order.size() * 4);
      p_rank.reserve((size_t)nfar * 4);
      const size_t per_slot = (size_t)(sc + sq) * esz + (size_t)sy * 36;
      int ns = g_cus;
      while (ns > 1 && (size_t)ns * per_slot > pool_budget()) ns /= 2;
      p_poolC.reserve((size_t)ns * sc * esz);
      p_poolQ.reserve((size_t)ns * sq * esz);
      p_Y.reserve((size_t)ns * sy * 36);
      p_slotC = sc, p_slotQ = sq, p_slotY = sy, p_nslots = ns, p_maxiter = MI, p_f64 = (int)f64;
      p_ready = true;
    }
    PAcaLaunch L{};
    L.prec = prec;
    L.xp = xp.as<float4>();
    L.fp = fp.as<float4>();
    L.u = u.as<double>();
    L.blk = p_blk.as<FarBlock>();
    L.order = p_order.as<int32_t>();
    L.rank_out = p_rank.as<int32_t>();
    L.nblocks = nfar;
    L.max_iter = MI;
    L.eps2 = epsilon * epsilon;
    L.use_symmetry = use_symmetry;
    L.poolC = p_poolC.p;
    L.poolQ = p_poolQ.p;
    L.slotC = p_slotC;
    L.slotQ = p_slotQ;
    L.Y = p_Y.as<float>();
    L.slotY = p_slotY;
    L.nslots = p_nslots;
    L.wf = devw_f(false);
    L.wh = devw_h(prec);
    HIPCHK(launch_persistent_aca(L, st));
    rank_src = p_rank.as<int32_t>();
    rank_pending = true;
  }

  bool rank_pending = false;
  const int32_t *rank_src = nullptr;  // device ranks of the last far field (far-list order)
  void fetch_ranks() {
    if (!rank_pending) return;
    HIPCHK(hipMemcpy(far_rank.data(), rank_src, far_rank.size() * 4, hipMemcpyDeviceToHost));
    rank_pending = false;
  }

  // ---- far field: batched ACA -------------------------------------------
  // Low-rank apply of the stopped blocks `done` (indices into hb), chunked
  // over I and J like the steps.
  size_t pool_budget() {
    if (far_pool_bytes == 0) {
      size_t fr = 0, tot = 0;
      HIPCHK(hipMemGetInfo(&fr, &tot));
      const size_t keep = (size_t)32 << 30;
      far_pool_bytes = std::min<size_t>((size_t)(0.7 * (double)fr), fr > keep ? fr - keep : 0);
      far_pool_bytes = std::max<size_t>(far_pool_bytes, (size_t)1 << 30);
    }
    return far_pool_bytes;
  }

  // The surrogate judge (hh_model_set_surrogate): per far block, the full
  // FP32 MLP against the surrogate on sampled entries; jd_flag[b] = 1 for the
  // blocks whose ACA takes surrogate entries.  The flags stay on the device;
  // fetch_judge() reads them (and the entry counts) when asked for.
  void judge_surrogate(hipStream_t st) {
    const int64_t nfar = (int64_t)far.size();
    judge_pending = false;
    sur_flag.assign((size_t)nfar, 0);
    sur_err.assign((size_t)nfar, 0.0f);
    for (auto &v : sur_stats) v = 0;
    if (!sur || nfar == 0) return;
    const int S = sur_samples;
    JudgeLaunch L{};
    L.blk = fb_all.as<FarBlock>();  // only is, m, js, n are read
    L.nblk = nfar;
    L.xp = xp.as<float4>();
    L.samples = S;
    L.r = jd_r.as<float>();
    L.m9 = jd_m9.as<float>();
    L.s6 = jd_s6.as<float>();
    L.params = sur->params.as<float>();
    L.r_lo = sur->r_lo;
    L.r_hi = sur->r_hi;
    L.tol = sur_tol;
    L.flag = jd_flag.as<int32_t>();
    L.err = jd_err.as<float>();
    L.wf = devw_f(false);
    L.wh = devw_h(prec);
    L.cus = g_cus;
    HIPCHK(launch_judge(L, st));
    judge_pending = true;
  }

  // Host copies of the last judged far field: flags, errors, entry counts.
  void fetch_judge() {
    if (!judge_pending) return;
    const int64_t nfar = (int64_t)far.size();
    HIPCHK(hipMemcpy(sur_flag.data(), jd_flag.p, (size_t)nfar * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sur_err.data(), jd_err.p, (size_t)nfar * 4, hipMemcpyDeviceToHost));
    int64_t ent[2] = {0, 0};
    HIPCHK(hipMemcpy(ent, ent_acc.p, 16, hipMemcpyDeviceToHost));
    sur_stats[0] = nfar;
    sur_stats[1] = sur_stats[2] = 0;
    for (int32_t f : sur_flag) sur_stats[f ? 1 : 2]++;
    sur_stats[3] = ent[1];
    sur_stats[4] = ent[0];
    judge_pending = false;
  }

  // Batches of far blocks whose C/Q pools fit the byte budget, their blocks
  // with pool offsets (one device array, far-list order), and every buffer
  // the iterations use sized for the largest batch.  Laid out once per tree
  // build and configuration.
  uint64_t far_key() const {
    uint64_t h = 1469598103934665603ull;
    for (uint64_t v : {(uint64_t)builds, (uint64_t)max_iter, (uint64_t)prec, (uint64_t)(sur != nullptr),
                       (uint64_t)far_c16, (uint64_t)far_q16, (uint64_t)aca_chunk, (uint64_t)far.size(),
                       (uint64_t)far_pool_bytes, (uint64_t)sur_samples}) {
      h ^= v;
      h *= 1099511628211ull;
    }
    return h;
  }

  void prepare_far_batches() {
    const int64_t nfar = (int64_t)far.size();
    (void)pool_budget();
    if (!batches.empty() && batch_key == far_key()) return;
    batches.clear();
    const bool f64 = prec == PREC_FP32;
    const bool c16 = pool16(true), q16 = pool16(false);
    const size_t escc = f64 ? 72 : (c16 ? 18 : 36), escq = f64 ? 72 : (q16 ? 18 : 36);
    const int MI = max_iter;
    auto block_bytes = [&](const FarBlock &f) {
      const int64_t kmax = std::min<int64_t>(MI, std::min(f.m, f.n));
      return kmax * ((int64_t)f.m * (int64_t)escc + (int64_t)f.n * (int64_t)escq);
    };
    // BF16 pools keep items in groups of 64 (1,152 bytes, the kernels'
    // aligned layout); FP32 / FP64 pools interleave 9 entries per item
    auto pool_bytes = [](int64_t items, size_t esc, bool grouped) {
      items = std::max<int64_t>(items, 1);
      return grouped ? (size_t)((items + 63) / 64) * 1152 : (size_t)items * esc;
    };
    std::vector<FarBlock> hb((size_t)nfar);
    size_t cmax = 1, qmax = 1;
    int64_t nbmax = 1, tcmax = 1, trmax = 1, ccmax = 1, crmax = 1, ymax = 1;
    const int CH = aca_chunk;
    int64_t b0 = 0;
    while (b0 < nfar) {
      FarBatch B{b0, b0, 0, 0, 0, 0, 0, 0};
      int64_t bytes = 0, co = 0, qo = 0;
      while (B.b1 < nfar) {
        FarBlock f = far[(size_t)B.b1];
        if (B.b1 > b0 && bytes + block_bytes(f) > (int64_t)pool_budget()) break;
        bytes += block_bytes(f);
        const int64_t kmax = std::min<int64_t>(MI, std::min(f.m, f.n));
        f.coff = co;
        f.qoff = qo;
        co += kmax * f.m;
        qo += kmax * f.n;
        hb[(size_t)B.b1] = f;
        B.ntc += (f.m + 15) / 16;
        B.ntr += (f.n + 15) / 16;
        B.ncc += (f.m + CH - 1) / CH;
        B.ncr += (f.n + CH - 1) / CH;
        B.ysum_c += f.m;
        B.ysum_r += f.n;
        B.b1++;
      }
      if (B.ntc > INT32_MAX || B.ntr > INT32_MAX || B.ncc > INT32_MAX || B.ncr > INT32_MAX)
        throw HHError(HH_ERR_ARG, "far-field batch too large for 32-bit tile indices");
      cmax = std::max(cmax, pool_bytes(co, escc, c16));
      qmax = std::max(qmax, pool_bytes(qo, escq, q16));
      nbmax = std::max(nbmax, B.b1 - B.b0);
      tcmax = std::max(tcmax, B.ntc), trmax = std::max(trmax, B.ntr);
      ccmax = std::max(ccmax, B.ncc), crmax = std::max(crmax, B.ncr);
      ymax = std::max(ymax, std::max(B.ysum_c, B.ysum_r));
      batches.push_back(B);
      b0 = B.b1;
    }
    upload(fb_all, hb.data(), hb.size() * sizeof(FarBlock));
    Cpool.reserve(cmax);
    Qpool.reserve(qmax);
    Y.reserve((size_t)ymax * 9 * 4);
    const size_t nb = (size_t)nbmax;
    it.reserve(nb * 4);
    ipiv.reserve(nb * MI * 4);
    jpiv.reserve(nb * (MI + 1) * 4);
    mu2.reserve(nb * 8);
    nu2.reserve(nb * 8);
    stop.reserve(nb * 4);
    for (auto &a : act2) a.reserve(nb * 4);
    ybase_c.reserve(nb * 8);
    ybase_r.reserve(nb * 8);
    toff_c.reserve(nb * 4);
    toff_r.reserve(nb * 4);
    for (auto *b : {&cfirst_c, &cfirst_r, &a_first_i, &a_first_j}) b->reserve((nb + 1) * 4);
    applylist.reserve(nb * 4);
    invb.reserve(nb * 9 * 8);
    tiles_fc.reserve((size_t)tcmax * 24);
    tiles_sc.reserve((size_t)tcmax * 24);
    tiles_fr.reserve((size_t)trmax * 24);
    tiles_sr.reserve((size_t)trmax * 24);
    part_c.reserve((size_t)ccmax * kAcaPartBytes);
    part_r.reserve((size_t)crmax * kAcaPartBytes);
    x_c.reserve((size_t)ccmax * MI * 9 * 8);
    x_r.reserve((size_t)crmax * MI * 9 * 8);
    w_i.reserve((size_t)ccmax * MI * 3 * 8);
    w_j.reserve((size_t)crmax * MI * 3 * 8);
    for (auto &c : cnt2) c.reserve(sizeof(AcaCounts));
    plan_part.reserve((size_t)kPlanGroups * kPlanFields * 8);
    rank_dev.reserve((size_t)std::max<int64_t>(nfar, 1) * 4);
    ent_acc.reserve(16);
    const int S = sur_samples;
    jd_r.reserve((size_t)std::max<int64_t>(nfar, 1) * S * 12);
    jd_m9.reserve((size_t)std::max<int64_t>(nfar, 1) * S * 36);
    jd_s6.reserve((size_t)std::max<int64_t>(nfar, 1) * S * 24);
    jd_flag.reserve((size_t)std::max<int64_t>(nfar, 1) * 4);
    jd_err.reserve((size_t)std::max<int64_t>(nfar, 1) * 4);
    for (auto *b : {&h_stop, &h_act, &h_toff_c, &h_toff_r, &h_list}) b->reserve(nb * 4);
    for (auto *b : {&h_ybase_c, &h_ybase_r}) b->reserve(nb * 8);
    for (auto *b : {&h_cfirst_c, &h_cfirst_r, &h_afirst_i, &h_afirst_j}) b->reserve((nb + 1) * 4);
    h_cnt.reserve(sizeof(AcaCounts));
    h_cnt_apply.reserve(sizeof(AcaCounts));
    batch_key = far_key();
  }

  // Launch descriptors of one batch (counts from cnt, grids from the bounds).
  AcaStepLaunch step_launch_for(const FarBatch &B, const FarBlock *blk, const AcaCounts *cnt,
                                const int32_t *active, int64_t grid_cap) {
    AcaStepLaunch sl{};
    sl.blk = blk;
    sl.active = active;
    sl.ybase_c = ybase_c.as<int64_t>();
    sl.ybase_r = ybase_r.as<int64_t>();
    sl.Y = Y.as<float>();
    sl.it = it.as<int32_t>();
    sl.ipiv = ipiv.as<int32_t>();
    sl.jpiv = jpiv.as<int32_t>();
    sl.mu2 = mu2.as<double>();
    sl.nu2 = nu2.as<double>();
    sl.stop = stop.as<int32_t>();
    sl.max_iter = max_iter;
    sl.eps2 = epsilon * epsilon;
    sl.cfirst_c = cfirst_c.as<int32_t>();
    sl.cfirst_r = cfirst_r.as<int32_t>();
    sl.part_c = part_c.p;
    sl.part_r = part_r.p;
    sl.x_c = x_c.as<double>();
    sl.x_r = x_r.as<double>();
    sl.inv = invb.as<double>();
    sl.chunk = aca_chunk;
    sl.cnt = cnt;
    sl.nactive = B.b1 - B.b0;
    sl.nchunks_c = B.ncc;
    sl.nchunks_r = B.ncr;
    sl.grid_cap = grid_cap;
    sl.pool_f64 = prec == PREC_FP32;
    sl.pool_c16 = pool16(true);
    sl.pool_q16 = pool16(false);
    sl.C = Cpool.p;
    sl.Q = Qpool.p;
    return sl;
  }

  AcaApplyLaunch apply_launch_for(const FarBatch &B, const FarBlock *blk, const AcaCounts *cnt,
                                  int64_t grid_cap) {
    AcaApplyLaunch al{};
    al.blk = blk;
    al.list = applylist.as<int32_t>();
    al.it = it.as<int32_t>();
    al.fp = fp.as<float4>();
    al.u = u.as<double>();
    al.use_symmetry = use_symmetry;
    al.cnt = cnt;
    al.nlist = B.b1 - B.b0;
    al.grid_cap = grid_cap;
    al.pool_f64 = prec == PREC_FP32;
    al.pool_c16 = pool16(true);
    al.pool_q16 = pool16(false);
    al.C = Cpool.p;
    al.Q = Qpool.p;
    al.first_i = a_first_i.as<int32_t>();
    al.first_j = a_first_j.as<int32_t>();
    al.nchunks_i = B.ncc;
    al.nchunks_j = B.ncr;
    al.w_i = w_i.as<double>();
    al.w_j = w_j.as<double>();
    al.max_iter = max_iter;
    al.chunk = aca_chunk;
    return al;
  }

  // One ACA iteration's passes over the active slots of cnt: per pass the
  // tile lists, the full-MLP entries, the surrogate entries, the steps.
  // hc: the host's copy of the counts (host control: exact grids), or NULL
  // (device control: grids from the batch bounds, at most grid_cap).
  // select_kind: 1 (only blocks on surrogate entries: no full-MLP launch), 0
  // (only full-MLP blocks: no surrogate launch), -1 (both kinds).
  void aca_passes(const FarBatch &B, const FarBlock *blk, const int32_t *kind, const AcaCounts *cnt,
                  const AcaCounts *hc, const int32_t *active, int64_t grid_cap, hipStream_t st,
                  bool timed, int select_kind = -1) {
    AcaStepLaunch sl = step_launch_for(B, blk, cnt, active, grid_cap);
    if (hc) sl.nactive = hc->na, sl.nchunks_c = hc->ncc, sl.nchunks_r = hc->ncr;
    for (int pass = 0; pass < 2; pass++) {
      AcaExpandLaunch ex{};
      ex.blk = blk;
      ex.active = active;
      ex.it = it.as<int32_t>();
      ex.ipiv = ipiv.as<int32_t>();
      ex.jpiv = jpiv.as<int32_t>();
      ex.ybase = pass ? ybase_r.as<int64_t>() : ybase_c.as<int64_t>();
      ex.toff = pass ? toff_r.as<int32_t>() : toff_c.as<int32_t>();
      ex.kind = kind;
      ex.cnt = cnt;
      ex.grid = hc ? hc->na : B.b1 - B.b0;
      ex.max_iter = max_iter;
      ex.row_pass = pass;
      ex.tiles_full = pass ? tiles_fr.p : tiles_fc.p;
      ex.tiles_sur = pass ? tiles_sr.p : tiles_sc.p;
      HIPCHK(launch_aca_expand(ex, st));
      AcaMlpLaunch ml{};
      ml.prec = prec;
      ml.xp = xp.as<float4>();
      ml.tiles = ex.tiles_full;
      ml.dntiles = pass ? &cnt->ntr[0] : &cnt->ntc[0];
      ml.ntiles = hc ? (pass ? hc->ntr[0] : hc->ntc[0]) : (pass ? B.ntr : B.ntc);
      ml.row_pass = pass;
      ml.Y = Y.as<float>();
      ml.wf = devw_f(false);
      ml.wh = devw_h(prec);
      ml.max_blocks = 8 * g_cus;
      ml.cus = mlp_cus > 0 ? mlp_cus : g_cus;
      hipEvent_t e0 = nullptr, e1 = nullptr;
      const bool mlp = select_kind != 1;
      if (timed && mlp) {
        while (evpool.size() < ev_used + 2) {
          hipEvent_t e;
          HIPCHK(hipEventCreate(&e));
          evpool.push_back(e);
        }
        e0 = evpool[ev_used], e1 = evpool[ev_used + 1];
        ev_used += 2;
        HIPCHK(hipEventRecord(e0, st));
      }
      if (mlp) HIPCHK(launch_aca_mlp(ml, st));
      if (timed && mlp) HIPCHK(hipEventRecord(e1, st));
      if (sur && kind && select_kind != 0) {
        AcaSurLaunch su{};
        su.xp = xp.as<float4>();
        su.tiles = ex.tiles_sur;
        su.dntiles = pass ? &cnt->ntr[1] : &cnt->ntc[1];
        su.ntiles = hc ? (pass ? hc->ntr[1] : hc->ntc[1]) : (pass ? B.ntr : B.ntc);
        su.row_pass = pass;
        su.Y = Y.as<float>();
        su.params = sur->params.as<float>();
        su.inv_scale = 1.0f / sur->r_hi;
        su.cus = g_cus;
        HIPCHK(launch_aca_sur(su, st));
      }
      HIPCHK(launch_aca_steps(sl, pass, st));
    }
  }
  size_t ev_used = 0;

  // Host control of one batch: per iteration the host builds the active
  // list and its prefix counts, uploads them, and reads the stop flags back.
  void batch_host_control(const FarBatch &B, const FarBlock *blk, const std::vector<FarBlock> &hb,
                          const int32_t *kind, const std::vector<int32_t> &hkind, hipStream_t st) {
    const int nb = (int)(B.b1 - B.b0);
    const int CH = aca_chunk;
    AcaCounts *hc = h_cnt.as<AcaCounts>(), *ha = h_cnt_apply.as<AcaCounts>();
    AcaCounts *dc = cnt2[0].as<AcaCounts>();
    std::vector<int32_t> act((size_t)nb);
    for (int b = 0; b < nb; b++) act[(size_t)b] = b;
    int64_t ent[2] = {0, 0};
    while (!act.empty()) {
      const int na = (int)act.size();
      int32_t *a_ = h_act.as<int32_t>(), *tc = h_toff_c.as<int32_t>(), *tr = h_toff_r.as<int32_t>();
      int32_t *fc = h_cfirst_c.as<int32_t>(), *fr = h_cfirst_r.as<int32_t>();
      int64_t *yc = h_ybase_c.as<int64_t>(), *yr = h_ybase_r.as<int64_t>();
      AcaCounts c{};
      for (int a = 0; a < na; a++) {
        const int b = act[(size_t)a];
        const FarBlock &f = hb[(size_t)b];
        const int kd = hkind.empty() ? 0 : hkind[(size_t)b];
        a_[a] = b;
        yc[a] = c.ysum_c;
        yr[a] = c.ysum_r;
        tc[a] = c.ntc[kd];
        tr[a] = c.ntr[kd];
        fc[a] = c.ncc;
        fr[a] = c.ncr;
        c.ysum_c += f.m;
        c.ysum_r += f.n;
        c.ntc[kd] += (f.m + 15) / 16;
        c.ntr[kd] += (f.n + 15) / 16;
        c.ncc += (f.m + CH - 1) / CH;
        c.ncr += (f.n + CH - 1) / CH;
        ent[kd] += f.m + f.n;
      }
      c.na = na;
      fc[na] = c.ncc;
      fr[na] = c.ncr;
      *hc = c;
      HIPCHK(hipMemcpyAsync(act2[0].p, a_, (size_t)na * 4, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(ybase_c.p, yc, (size_t)na * 8, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(ybase_r.p, yr, (size_t)na * 8, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(toff_c.p, tc, (size_t)na * 4, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(toff_r.p, tr, (size_t)na * 4, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(cfirst_c.p, fc, (size_t)(na + 1) * 4, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(cfirst_r.p, fr, (size_t)(na + 1) * 4, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(dc, hc, sizeof(AcaCounts), hipMemcpyHostToDevice, st));
      aca_passes(B, blk, kind, dc, &c, act2[0].as<int32_t>(), INT64_MAX, st, true);
      HIPCHK(hipMemcpyAsync(h_stop.p, stop.p, (size_t)na * 4, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      const int32_t *hs = h_stop.as<int32_t>();
      std::vector<int32_t> done, keep;
      for (int a = 0; a < na; a++) (hs[a] ? done : keep).push_back(act[(size_t)a]);
      if (!done.empty()) {
        const int nl = (int)done.size();
        int32_t *hl = h_list.as<int32_t>(), *fi = h_afirst_i.as<int32_t>(), *fj = h_afirst_j.as<int32_t>();
        AcaCounts d = c;
        d.nl = nl;
        d.nci = d.ncj = 0;
        for (int a = 0; a < nl; a++) {
          const FarBlock &f = hb[(size_t)done[(size_t)a]];
          hl[a] = done[(size_t)a];
          fi[a] = d.nci;
          fj[a] = d.ncj;
          d.nci += (f.m + CH - 1) / CH;
          d.ncj += (f.n + CH - 1) / CH;
        }
        fi[nl] = d.nci;
        fj[nl] = d.ncj;
        *ha = d;
        HIPCHK(hipMemcpyAsync(applylist.p, hl, (size_t)nl * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(a_first_i.p, fi, (size_t)(nl + 1) * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(a_first_j.p, fj, (size_t)(nl + 1) * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dc, ha, sizeof(AcaCounts), hipMemcpyHostToDevice, st));
        AcaApplyLaunch al = apply_launch_for(B, blk, dc, INT64_MAX);
        al.nlist = nl;
        al.nchunks_i = d.nci;
        al.nchunks_j = d.ncj;
        HIPCHK(launch_aca_apply(al, st));
        HIPCHK(hipStreamSynchronize(st));  // the pinned lists are rewritten next iteration
      }
      act.swap(keep);
    }
    // entry counts of this batch onto the device accumulator
    int64_t acc[2] = {0, 0};
    HIPCHK(hipMemcpy(acc, ent_acc.p, 16, hipMemcpyDeviceToHost));
    acc[0] += ent[0];
    acc[1] += ent[1];
    HIPCHK(hipMemcpy(ent_acc.p, acc, 16, hipMemcpyHostToDevice));
  }

  // Device control of one batch: the first plan takes every block; each
  // iteration's finalize step writes the stop flags, the next plan compacts
  // the active list and lists the stopped blocks, which are applied at once.
  // max_iter iterations stop every block.  With early_exit (not under stream
  // capture) each plan's active count is copied to pinned memory; the host
  // waits for the copy kLag iterations back -- the device still has kLag
  // iterations queued -- and stops enqueuing once that count is zero.
  static constexpr int kLag = 3, kRing = 8;
  PBuf h_na;
  hipEvent_t ev_na[kRing] = {};
  void batch_device_control(const FarBatch &B, const FarBlock *blk, const int32_t *kind,
                            hipStream_t st, bool timed, bool early_exit, int select_kind) {
    if (early_exit && !ev_na[0]) {
      for (auto &e : ev_na) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      h_na.reserve(kRing * sizeof(int32_t));
    }
    const int nb = (int)(B.b1 - B.b0);
    const int64_t cap = 32 * (int64_t)g_cus;
    AcaPlanLaunch pl{};
    pl.blk = blk;
    pl.kind = kind;
    pl.nb = nb;
    pl.chunk = aca_chunk;
    pl.ybase_c = ybase_c.as<int64_t>();
    pl.ybase_r = ybase_r.as<int64_t>();
    pl.toff_c = toff_c.as<int32_t>();
    pl.toff_r = toff_r.as<int32_t>();
    pl.cfirst_c = cfirst_c.as<int32_t>();
    pl.cfirst_r = cfirst_r.as<int32_t>();
    pl.alist = applylist.as<int32_t>();
    pl.first_i = a_first_i.as<int32_t>();
    pl.first_j = a_first_j.as<int32_t>();
    pl.part = plan_part.as<int64_t>();
    pl.select_kind = kind ? select_kind : -1;
    pl.act_in = nullptr, pl.stop_in = nullptr, pl.cnt_in = nullptr;
    pl.act_out = act2[0].as<int32_t>();
    pl.cnt_out = cnt2[0].as<AcaCounts>();
    HIPCHK(launch_aca_plan(pl, st));
    int cur = 0;
    for (int iter = 0; iter < max_iter; iter++) {
      const AcaCounts *c = cnt2[cur].as<AcaCounts>();
      aca_passes(B, blk, kind, c, nullptr, act2[cur].as<int32_t>(), cap, st, timed,
                 kind ? select_kind : -1);
      const int nxt = cur ^ 1;
      pl.act_in = act2[cur].as<int32_t>();
      pl.stop_in = stop.as<int32_t>();
      pl.cnt_in = c;
      pl.act_out = act2[nxt].as<int32_t>();
      pl.cnt_out = cnt2[nxt].as<AcaCounts>();
      HIPCHK(launch_aca_plan(pl, st));
      HIPCHK(launch_aca_apply(apply_launch_for(B, blk, cnt2[nxt].as<AcaCounts>(), cap), st));
      cur = nxt;
      if (early_exit) {
        int32_t *na = h_na.as<int32_t>();
        HIPCHK(hipMemcpyAsync(na + iter % kRing, &cnt2[cur].as<AcaCounts>()->na, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(ev_na[iter % kRing], st));
        if (iter >= kLag) {
          const int back = (iter - kLag) % kRing;
          HIPCHK(hipEventSynchronize(ev_na[back]));
          if (na[back] == 0) break;  // every block stopped and was applied by then
        }
      }
    }
    HIPCHK(launch_aca_entries(cnt2[cur].as<AcaCounts>(), ent_acc.as<int64_t>(), st));
  }

  // Batched far field: judge, then per batch the iterations under device or
  // host control; ranks into rank_dev (read when asked for).  With a
  // surrogate under device control the iterations can run as two loops per
  // batch (far_loops with select_kind 1, then 0): blocks on surrogate entries
  // need no LDS-resident MLP kernel and run beside the near field; the
  // judge-rejected blocks follow once the near field is done.
  bool split_far() const { return sur && aca_control == 1; }

  void far_setup(hipStream_t st) {
    prepare_far_batches();
    judge_surrogate(st);
    HIPCHK(launch_zero_f64(ent_acc.as<double>(), 2, st));  // two int64 zeros
    ev_used = 0;
  }

  void far_loops(hipStream_t st, bool timed, int select_kind) {
    const int64_t nfar = (int64_t)far.size();
    const bool host_ctl = aca_control == 0;
    std::vector<int32_t> hkind;
    std::vector<FarBlock> hb;
    if (host_ctl) {
      if (sur) {
        fetch_judge();
        hkind = sur_flag;
        judge_pending = true;  // entry counts still to come
      }
      hb.resize((size_t)nfar);
      HIPCHK(hipMemcpy(hb.data(), fb_all.p, (size_t)nfar * sizeof(FarBlock), hipMemcpyDeviceToHost));
    }
    for (const FarBatch &B : batches) {
      const int nb = (int)(B.b1 - B.b0);
      if (verbose) {
        std::fprintf(stderr, "[hh] far blocks %lld..%lld of %lld (%s control, kind %d)\n",
                     (long long)B.b0, (long long)B.b1, (long long)nfar, host_ctl ? "host" : "device",
                     select_kind);
        std::fflush(stderr);
      }
      const FarBlock *blk = fb_all.as<FarBlock>() + B.b0;
      const int32_t *kind = sur ? jd_flag.as<int32_t>() + B.b0 : nullptr;
      HIPCHK(launch_aca_init(blk, nb, max_iter, it.as<int32_t>(), jpiv.as<int32_t>(), mu2.as<double>(),
                             nu2.as<double>(), st));
      if (host_ctl) {
        std::vector<FarBlock> hbb(hb.begin() + B.b0, hb.begin() + B.b1);
        std::vector<int32_t> hk;
        if (!hkind.empty()) hk.assign(hkind.begin() + B.b0, hkind.begin() + B.b1);
        batch_host_control(B, blk, hbb, kind, hk, st);
      } else {
        batch_device_control(B, blk, kind, st, timed, timed, select_kind);
      }
      HIPCHK(launch_copy_ranks(it.as<int32_t>(), nb, rank_dev.as<int32_t>() + B.b0,
                               kind && select_kind >= 0 ? kind : nullptr, select_kind, st));
    }
  }

  void far_finish() {
    if (sur) judge_pending = true;
    rank_src = rank_dev.as<int32_t>();
    rank_pending = true;
  }

  void far_field(hipStream_t st, bool timed = true) {
    if (resolved_far_mode() == 0) {
      far_field_persistent(st);
      return;
    }
    if (far.empty()) return;
    far_setup(st);
    if (split_far()) {
      far_loops(st, timed, 1);
      far_loops(st, timed, 0);
    } else {
      far_loops(st, timed, -1);
    }
    far_finish();
  }

  // ACA pass times (full-MLP entries) of the last batched far field, after
  // the stream has completed it.
  void collect_aca_profile() {
    prof_ms[1] = 0;
    for (size_t i = 0; i + 1 < ev_used; i += 2) {
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, evpool[i], evpool[i + 1]));
      prof_ms[1] += t;
    }
    prof_launches[1] = (int64_t)(ev_used / 2);
    ev_used = 0;
    int64_t ent[2] = {0, 0};
    HIPCHK(hipMemcpy(ent, ent_acc.p, 16, hipMemcpyDeviceToHost));
    prof_pairs[1] = ent[0];
  }

  void near_field(hipStream_t st, const DBuf &blocks, const DBuf &tasks, int64_t ntasks,
                  int p, double *uacc, const int32_t *tlist) {
    NearLaunch L{};
    L.prec = p;
    L.hcfg = near_hcfg;
    L.xp = xp.as<float4>();
    L.fp = fp.as<float4>();
    L.tlist = tlist;
    L.blocks = blocks.as<RowBlock>();
    L.tasks = tasks.as<NearTask>();
    L.ntasks = ntasks;
    L.u = uacc;
    L.wf = devw_f(p == PREC_FP64);
    L.wh = devw_h(p);
    L.max_blocks = 8 * g_cus;
    L.cus = near_workgroups();
    // one near launch of a matvec at a time per model: the persistent 16-bit
    // kernel's work queue
    L.next = (p == PREC_BF16 || p == PREC_FP16) ? near_next.as<int32_t>() : nullptr;
    HIPCHK(launch_near(L, st));
  }

  void check_ready() {
    if (!has_weights) throw HHError(HH_ERR_STATE, "two-body model not loaded");
    if (!built) update();  // the reference would dereference empty lists here
    ensure_buffers();
    if (!ev[0]) {
      for (auto &e : ev) HIPCHK(hipEventCreate(&e));
      for (auto &e : evm) HIPCHK(hipEventCreate(&e));
      for (auto &e : ev_all) HIPCHK(hipEventCreate(&e));
      HIPCHK(hipEventCreate(&far0_ev));
      HIPCHK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming));
      HIPCHK(hipStreamCreate(&side));  // default priority
    }
  }

  // Near field on the side stream and the persistent far field on st, both
  // after the forces are gathered and u is cleared on st; st continues once
  // both are done.  Both accumulate into u with FP64 atomics.  With timed,
  // ev[0..1] bracket the near field and far0_ev / ev[2] the far field.
  void near_beside_persistent_far(hipStream_t st, bool timed) {
    near_beside_far(st, timed, true);
  }

  // Near field on the side stream, the far field (persistent kernel or
  // batched iterations) on st; st continues once both are done.  The
  // batched far field's per-iteration host waits are on st only, so the
  // near kernel keeps running through them.
  void near_beside_far(hipStream_t st, bool timed, bool persistent) {
    // the judge first (its FP32 MLP would only compete with the near field)
    const bool split = !persistent && split_far() && !far.empty();
    if (timed) HIPCHK(hipEventRecord(far0_ev, st));
    if (split) far_setup(st);
    HIPCHK(hipEventRecord(fork_ev, st));
    HIPCHK(hipStreamWaitEvent(side, fork_ev, 0));
    if (timed) HIPCHK(hipEventRecord(ev[0], side));
    near_field(side, near_blocks, near_tasks, n_near_tasks, prec, u.as<double>(), nullptr);
    if (timed) HIPCHK(hipEventRecord(ev[1], side));
    HIPCHK(hipEventRecord(join_ev, side));
    // The judge-rejected blocks' loop is latency-bound (few blocks, run to
    // max_iter) and its full-MLP kernel needs a whole compute unit's LDS.
    // When the 16-bit near kernel leaves compute units free, that loop runs
    // first, beside the near kernel, and the surrogate blocks' loop follows;
    // otherwise it waits for the near kernel to finish.
    const bool rejected_first =
        split && (prec == PREC_BF16 || prec == PREC_FP16) && near_workgroups() < g_cus;
    if (persistent) {
      far_field_persistent(st);
    } else if (split) {
      if (rejected_first) {
        // one full-MLP workgroup per compute unit the near kernel leaves
        // free: the kernel's workgroups split its tiles evenly, so more than
        // fit at once would run in two rounds
        mlp_cus = g_cus - near_workgroups();
        far_loops(st, timed, 0);
        mlp_cus = 0;
      }
      far_loops(st, timed, 1);  // beside the near field
    } else {
      far_field(st, timed);
    }
    if (timed && !split) HIPCHK(hipEventRecord(ev[2], st));
    HIPCHK(hipStreamWaitEvent(st, join_ev, 0));
    if (split) {
      if (!rejected_first) far_loops(st, timed, 0);  // full MLP, after the near field
      far_finish();
      if (timed) HIPCHK(hipEventRecord(ev[2], st));
    }
  }

  int resolved_far_mode() const {
    if (sur) return 1;  // surrogate entries: batched iterations only
    return far_mode >= 0 ? far_mode : ((int64_t)far.size() <= (int64_t)g_cus ? 0 : 1);
  }

  // The matvec can be enqueued inside a stream capture: a kept tree (no host
  // build), the far field on the persistent kernel (no host reads per ACA
  // iteration) and its buffers set up by an earlier dot.
  bool capturable() const {
    if (!built || !tree_reuse || !has_weights || !ev[0]) return false;
    if (fp.bytes < 16 * (size_t)n || u.bytes < 24 * (size_t)n) return false;
    if (far.empty()) return true;
    if (resolved_far_mode() == 0)
      return p_ready && p_maxiter == max_iter && p_f64 == (int)(prec == PREC_FP32);
    // batched iterations under device control, laid out by an earlier dot
    return aca_control == 1 && !batches.empty() && batch_key == far_key();
  }

  // Changes whenever something a captured matvec reads or launches with does.
  uint64_t graph_signature() const {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) {
      h ^= v;
      h *= 1099511628211ull;
    };
    uint64_t e;
    std::memcpy(&e, &epsilon, 8);
    uint64_t tol;
    std::memcpy(&tol, &sur_tol, 8);
    for (uint64_t v : {(uint64_t)prec, (uint64_t)max_iter, e, (uint64_t)use_symmetry,
                       (uint64_t)n_near_tasks, (uint64_t)p_nslots, (uint64_t)p_slotC,
                       (uint64_t)p_slotQ, (uint64_t)p_slotY, (uint64_t)far.size(), builds, loads,
                       (uint64_t)resolved_far_mode(), (uint64_t)aca_control, batch_key,
                       (uint64_t)(uintptr_t)sur, tol, (uint64_t)sur_samples, (uint64_t)near_overlap,
                       (uint64_t)near_hcfg, (uint64_t)near_cus})
      mix(v);
    for (const DBuf *b : {&xp, &perm, &fp, &u, &near_blocks, &near_tasks, &p_blk, &p_order, &p_rank,
                          &p_poolC, &p_poolQ, &p_Y, &fb_all, &Cpool, &Qpool, &Y, &x_c, &x_r,
                          &part_c, &part_r, &w_i, &w_j, &jd_flag, &rank_dev, &near_next})
      mix((uint64_t)(uintptr_t)b->p);
    for (const DBuf *arr : {wf32, wf64}) for (int k = 0; k < 5; k++) mix((uint64_t)(uintptr_t)arr[k].p);
    for (int k = 0; k < 3; k++) mix((uint64_t)(uintptr_t)wh_f[k].p);
    for (int k = 0; k < 2; k++) mix((uint64_t)(uintptr_t)wh_bf16[k].p), mix((uint64_t)(uintptr_t)wh_f16[k].p);
    return h;
  }

  // Far-field launch tuning on the caller's forces: each candidate (the
  // persistent kernel when there are at most 4 far blocks per compute unit;
  // the batched iterations with step chunks of 64, 128 and 256 items) runs
  // one matvec untimed and one timed (HIP events around the whole matvec,
  // since the near field runs beside the far field); the fastest is kept.
  // Rows of `out`: {mode, chunk, matvec ms}.
  void tune_far_field(const float *f_dev, hipStream_t st, std::vector<std::array<double, 3>> &out) {
    check_ready();
    std::vector<std::array<int, 2>> cand;
    if ((int64_t)far.size() <= 4 * (int64_t)g_cus) cand.push_back({0, aca_chunk});
    for (int c : {64, 128, 256})
      if (c <= kAcaChunk) cand.push_back({1, c});
    out.clear();
    double best = 0;
    std::array<int, 2> pick = cand.front();
    const int mode0 = far_mode, chunk0 = aca_chunk;
    try {
      for (const auto &c : cand) {
        far_mode = c[0];
        aca_chunk = c[1];
        for (int rep = 0; rep < 2; rep++) dot_device(uout.as<float>(), f_dev, st);
        HIPCHK(hipStreamSynchronize(st));
        // the whole matvec: the near field may run beside the far field
        out.push_back({(double)c[0], (double)c[1], t_total});
        if (out.size() == 1 || t_total < best) best = t_total, pick = c;
      }
    } catch (...) {  // a failed candidate leaves the settings as they were
      far_mode = mode0;
      aca_chunk = chunk0;
      throw;
    }
    far_mode = pick[0];
    aca_chunk = pick[1];
  }

  int near_workgroups() const { return near_cus > 0 && near_cus < g_cus ? near_cus : g_cus; }

  // Launch tuning of the 16-bit near kernel on the caller's forces: each
  // candidate runs one untimed and one timed matvec (HIP events around the
  // whole matvec: the near field runs beside the far field, so the best
  // choice depends on both); the fastest is kept.  Candidates: every tile
  // shape on one workgroup per compute unit, and, when the far field runs
  // beside the near field, the first three shapes on 5/8, 1/2 and 3/8 of the
  // compute units (the far kernels then hold the other compute units at full
  // occupancy instead of the one wave per SIMD that fits beside a near
  // workgroup).  Rows of `out`: {shape, tiles, threads, workgroups, matvec ms}.
  void tune_near(const float *f_dev, hipStream_t st, std::vector<std::array<double, 5>> &out) {
    check_ready();
    out.clear();
    if (prec != PREC_BF16 && prec != PREC_FP16) return;  // the FP32 kernel has one shape
    std::vector<std::array<int, 2>> cand;
    for (int c = 0; c < kNearShapeCount; c++) cand.push_back({c, 0});
    if (!far.empty() && (resolved_far_mode() == 0 || near_overlap))
      for (int wg : {g_cus * 5 / 8, g_cus / 2, g_cus * 3 / 8})
        for (int c = 0; c < 3; c++) cand.push_back({c, wg});
    const int h0 = near_hcfg, w0 = near_cus;
    std::array<int, 2> pick{h0, w0};
    double best = 0;
    try {
      for (const auto &c : cand) {
        near_hcfg = c[0];
        near_cus = c[1];
        for (int rep = 0; rep < 2; rep++) dot_device(uout.as<float>(), f_dev, st);
        HIPCHK(hipStreamSynchronize(st));
        out.push_back({(double)c[0], (double)kNearShapes[c[0]].tiles,
                       (double)kNearShapes[c[0]].threads, (double)near_workgroups(), t_total});
        if (out.size() == 1 || t_total < best) best = t_total, pick = c;
      }
    } catch (...) {
      near_hcfg = h0;
      near_cus = w0;
      throw;
    }
    near_hcfg = pick[0];
    near_cus = pick[1];
  }

  // u = M(x) f without host synchronisation or timing (capturable state only).
  void enqueue_dot(float *u_dev, const float *f_dev, hipStream_t st) {
    if (!capturable()) throw HHError(HH_ERR_STATE, "matvec not capturable in the current state");
    HIPCHK(launch_gather_forces(f_dev, n, perm.as<int32_t>(), fp.as<float4>(), st));
    HIPCHK(launch_zero_f64(u.as<double>(), 3 * n, st));
    const bool persistent = resolved_far_mode() == 0;
    if (!far.empty() && (persistent || near_overlap)) {
      near_beside_far(st, false, persistent);
    } else {
      near_field(st, near_blocks, near_tasks, n_near_tasks, prec, u.as<double>(), nullptr);
      if (!far.empty()) far_field(st, false);
    }
    HIPCHK(launch_scatter_velocity(u.as<double>(), n, perm.as<int32_t>(), u_dev, st));
  }

  void dot_device(float *u_dev, const float *f_dev, hipStream_t st) {
    check_ready();
    HIPCHK(hipEventRecord(ev_all[0], st));
    HIPCHK(launch_gather_forces(f_dev, n, perm.as<int32_t>(), fp.as<float4>(), st));
    HIPCHK(hipMemsetAsync(u.p, 0, 24 * (size_t)n, st));
    prof_ms[1] = 0;
    prof_pairs[1] = prof_launches[1] = 0;
    const bool persistent = resolved_far_mode() == 0;
    const bool beside = !far.empty() && (persistent || near_overlap);
    if (beside) {
      near_beside_far(st, true, persistent);
    } else {
      HIPCHK(hipEventRecord(ev[0], st));
      near_field(st, near_blocks, near_tasks, n_near_tasks, prec, u.as<double>(), nullptr);
      HIPCHK(hipEventRecord(ev[1], st));
      HIPCHK(hipEventRecord(far0_ev, st));
      far_field(st);
      HIPCHK(hipEventRecord(ev[2], st));
    }
    HIPCHK(launch_scatter_velocity(u.as<double>(), n, perm.as<int32_t>(), u_dev, st));
    HIPCHK(hipEventRecord(ev_all[1], st));
    HIPCHK(hipEventSynchronize(ev_all[1]));
    if (!persistent && !far.empty()) collect_aca_profile();
    float a = 0, b = 0, c = 0;
    HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
    HIPCHK(hipEventElapsedTime(&b, far0_ev, ev[2]));
    HIPCHK(hipEventElapsedTime(&c, ev_all[0], ev_all[1]));
    t_close = a;
    t_far = b;
    t_total = c;  // the whole matvec, gather to scatter (near and far may overlap)
    prof_ms[0] = a;  // the near field is one kernel launch
    prof_pairs[0] = near_pairs;
    prof_launches[0] = n_near_tasks > 0 ? 1 : 0;
  }

  void dot_host(float *uh, const float *fh, bool dense) {
    check_ready();
    HIPCHK(hipMemcpy(fin.p, fh, 12 * (size_t)n, hipMemcpyHostToDevice));
    if (!dense) {
      dot_device(uout.as<float>(), fin.as<float>(), stream);
    } else {
      build_dense();
      HIPCHK(launch_gather_forces(fin.as<float>(), n, perm.as<int32_t>(), fp.as<float4>(), stream));
      HIPCHK(hipMemsetAsync(u.p, 0, 24 * (size_t)n, stream));
      near_field(stream, dense_blocks, dense_tasks, n_dense_tasks, PREC_FP32, u.as<double>(),
                 nullptr);
      HIPCHK(launch_scatter_velocity(u.as<double>(), n, perm.as<int32_t>(), uout.as<float>(),
                                     stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(uh, uout.p, 12 * (size_t)n, hipMemcpyDeviceToHost));
  }

  // DenseDot blocks: every leaf x leaf pair, symmetrised (src/DenseDot.cpp).
  void build_dense() {
    if (n_dense_tasks >= 0) return;
    std::vector<RowBlock> rb;
    std::vector<NearTask> tk;
    for (int64_t I : tl.leaves)
      for (int64_t J : tl.leaves) {
        RowBlock r{(int32_t)tl.start(I), (int32_t)tl.size(I), (int32_t)tl.start(J),
                   (int32_t)tl.size(J), BF_SYMM, 0, 0, 0};
        const int32_t b = (int32_t)rb.size();
        rb.push_back(r);
        for (int32_t o = 0; o < r.m; o += 4) tk.push_back(NearTask{b, o});
      }
    upload(dense_blocks, rb.data(), rb.size() * sizeof(RowBlock));
    upload(dense_tasks, tk.data(), tk.size() * sizeof(NearTask));
    n_dense_tasks = (int64_t)tk.size();
  }

  // FP64 evaluation at selected targets (PostCheckDot).
  void check_fp64(const float *fh, const int64_t *targets, int64_t nt, int conv, double *out) {
    check_ready();
    HIPCHK(hipMemcpy(fin.p, fh, 12 * (size_t)n, hipMemcpyHostToDevice));
    HIPCHK(launch_gather_forces(fin.as<float>(), n, perm.as<int32_t>(), fp.as<float4>(), stream));
    std::vector<int64_t> inv((size_t)n);
    for (int64_t i = 0; i < n; i++) inv[(size_t)tl.perm[(size_t)i]] = i;
    // leaf of every tree position, parent of every node
    std::vector<int64_t> pos_leaf((size_t)n, -1), parent((size_t)tl.T, -1);
    for (int64_t lf : tl.leaves)
      for (int64_t p = tl.start(lf); p < tl.start(lf) + tl.size(lf); p++) pos_leaf[(size_t)p] = lf;
    for (int64_t nd = 0; nd < tl.T; nd++)
      if (!tl.leaf(nd)) {
        parent[(size_t)tl.rows[5 * nd]] = nd;
        parent[(size_t)tl.rows[5 * nd + 1]] = nd;
      }
    // targets grouped by leaf; tlist holds tree positions
    std::vector<std::pair<int64_t, int64_t>> order;  // (leaf, sample index)
    for (int64_t q = 0; q < nt; q++) {
      if (targets[q] < 0 || targets[q] >= n) throw HHError(HH_ERR_ARG, "target out of range");
      const int64_t p = inv[(size_t)targets[q]];
      order.push_back({pos_leaf[(size_t)p], q});
    }
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<int64_t, int64_t> &a, const std::pair<int64_t, int64_t> &b) {
                       return a.first < b.first;
                     });
    std::vector<int32_t> tlist((size_t)nt);
    for (int64_t q = 0; q < nt; q++) tlist[(size_t)q] = (int32_t)inv[(size_t)targets[order[(size_t)q].second]];
    // per-node source lists of the block convention
    struct Src {
      int32_t j0, n, flags;
    };
    std::vector<std::vector<Src>> per((size_t)tl.T);
    if (conv == 0) {
      auto add = [&](const std::vector<int64_t> &BI, const std::vector<int64_t> &BJ, bool isfar) {
        for (size_t b = 0; b < BI.size(); b++) {
          const int64_t I = BI[b], J = BJ[b];
          const int32_t fl = isfar ? BF_SYMM : 0;
          per[(size_t)I].push_back({(int32_t)tl.start(J), (int32_t)tl.size(J), fl});
          if (use_symmetry && J != I)
            per[(size_t)J].push_back({(int32_t)tl.start(I), (int32_t)tl.size(I), fl | BF_NEG});
        }
      };
      add(tl.closeI, tl.closeJ, false);
      add(tl.farI, tl.farJ, true);
    }
    std::vector<RowBlock> rb;
    std::vector<NearTask> tk;
    int64_t q = 0;
    while (q < nt) {
      const int64_t lf = order[(size_t)q].first;
      int64_t q1 = q;
      while (q1 < nt && order[(size_t)q1].first == lf) q1++;
      const int32_t m = (int32_t)(q1 - q);
      std::vector<Src> srcs;
      if (conv == 0) {
        for (int64_t nd = lf; nd >= 0; nd = parent[(size_t)nd])
          for (const Src &s : per[(size_t)nd]) srcs.push_back(s);
      } else {
        srcs.push_back({0, (int32_t)n, BF_SYMM});
      }
      for (const Src &s : srcs) {
        const int32_t b = (int32_t)rb.size();
        rb.push_back(RowBlock{(int32_t)q, m, s.j0, s.n, s.flags | BF_EXPLICIT_I, 0, 0, 0});
        for (int32_t o = 0; o < m; o += 4) tk.push_back(NearTask{b, o});
      }
      q = q1;
    }
    upload(chk_blocks, rb.data(), rb.size() * sizeof(RowBlock));
    upload(chk_tasks, tk.data(), tk.size() * sizeof(NearTask));
    upload(chk_tlist, tlist.data(), tlist.size() * 4);
    chk_u.reserve(24 * (size_t)std::max<int64_t>(nt, 1));
    HIPCHK(hipMemsetAsync(chk_u.p, 0, 24 * (size_t)nt, stream));
    near_field(stream, chk_blocks, chk_tasks, (int64_t)tk.size(), PREC_FP64, chk_u.as<double>(),
               chk_tlist.as<int32_t>());
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<double> tmp(3 * (size_t)nt);
    HIPCHK(hipMemcpy(tmp.data(), chk_u.p, 24 * (size_t)nt, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < nt; k++) {
      const int64_t qq = order[(size_t)k].second;
      for (int d = 0; d < 3; d++) out[3 * qq + d] = tmp[3 * (size_t)k + d];
    }
  }
};

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static void export_lists(const TreeLists &t, uint64_t *tree5, uint64_t *perm, float *aux6,
                         int64_t *close_ij, int64_t *far_ij, int64_t *leaves);

extern "C" {

const char *hh_last_error(void) { return g_last_error.c_str(); }

int hh_init(void) {
  return guarded([&]() {
    ensure_device();
    (void)scratch_ptr();  // the no-allocation gradient's workspace address
  });
}

int hh_finalize(void) {
  return guarded([&]() {
    if (g_device >= 0) HIPCHK(hipDeviceSynchronize());
  });
}

int hh_model_create_instances(const float *coord, const int64_t *counts, int64_t ninst,
                              int block_size, hh_model **out) {
  return guarded([&]() {
    if (!out || !counts) throw HHError(HH_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (ninst <= 0) throw HHError(HH_ERR_ARG, "need at least one instance");
    int64_t n = 0;
    for (int64_t k = 0; k < ninst; k++) {
      if (counts[k] <= 0) throw HHError(HH_ERR_ARG, "every instance needs particles");
      n += counts[k];
    }
    hh_model *m = nullptr;
    const int rc = hh_model_create(coord, n, block_size, &m);
    if (rc != HH_OK) throw HHError(rc, g_last_error);
    m->inst_counts.assign(counts, counts + ninst);
    *out = m;
  });
}

int hh_model_create(const float *coord, int64_t n, int block_size, hh_model **out) {
  return guarded([&]() {
    if (!out) throw HHError(HH_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (n <= 0 || !coord) throw HHError(HH_ERR_ARG, "need n > 0 coordinates");
    if (n > (int64_t)INT32_MAX / 4) throw HHError(HH_ERR_ARG, "too many particles");
    if (block_size <= 0) throw HHError(HH_ERR_ARG, "block size must be positive");
    int chunk = kAcaChunk;
    if (const char *ch = std::getenv("HH_ACA_CHUNK")) {
      chunk = std::atoi(ch);
      if (chunk < 1 || chunk > kAcaChunk)
        throw HHError(HH_ERR_ARG, "HH_ACA_CHUNK must be in 1.." + std::to_string(kAcaChunk));
    }
    ensure_device();
    hh_model *m = new hh_model();
    m->n = n;
    m->leaf = block_size;
    m->x.assign(coord, coord + 3 * n);
    if (hipStreamCreate(&m->stream) != hipSuccess) {
      delete m;
      throw HHError(HH_ERR_DEVICE, "hipStreamCreate failed");
    }
    if (const char *fm = std::getenv("HH_FAR_MODE")) m->far_mode = std::atoi(fm);
    if (const char *tb = std::getenv("HH_TREE_BUILD")) m->tree_device = std::strcmp(tb, "host") != 0;
    if (const char *ns = std::getenv("HH_NEAR_SHAPE")) {
      const int v = std::atoi(ns);
      if (v < 0 || v >= kNearShapeCount) throw HHError(HH_ERR_ARG, "HH_NEAR_SHAPE out of range");
      m->near_hcfg = v;
    }
    if (const char *nw = std::getenv("HH_NEAR_WORKGROUPS")) {
      const int v = std::atoi(nw);
      if (v < 0) throw HHError(HH_ERR_ARG, "HH_NEAR_WORKGROUPS is negative");
      m->near_cus = v;
    }
    if (const char *c16 = std::getenv("HH_FAR_C16")) m->far_c16 = std::atoi(c16) != 0;
    if (const char *q16 = std::getenv("HH_FAR_Q16")) m->far_q16 = std::atoi(q16) != 0;
    if (const char *vb = std::getenv("HH_VERBOSE")) m->verbose = std::atoi(vb) != 0;
    if (const char *ov = std::getenv("HH_NEAR_OVERLAP")) m->near_overlap = std::atoi(ov) != 0;
    if (const char *ac = std::getenv("HH_ACA_CONTROL")) m->aca_control = std::strcmp(ac, "host") != 0;
    m->aca_chunk = chunk;
    if (const char *gb = std::getenv("HH_FAR_POOL_GB"))
      m->far_pool_bytes = (size_t)(std::atof(gb) * (double)(1ull << 30));
    *out = m;
  });
}

void hh_model_destroy(hh_model *m) { delete m; }

int hh_model_load_two_body(hh_model *m, const float *weights, int64_t count) {
  return guarded([&]() {
    if (!m || !weights) throw HHError(HH_ERR_ARG, "NULL argument");
    if (count != 71177)
      throw HHError(HH_ERR_ARG, "expected 71177 weights (3-128-512-9 MLP), got " + std::to_string(count));
    m->load_weights(weights);
  });
}

#define HH_SETTER(name, type, field, check)                                     \
  int hh_model_set_##name(hh_model *m, type v) {                               \
    return guarded([&]() {                                                      \
      if (!m) throw HHError(HH_ERR_ARG, "NULL model");                          \
      check;                                                                    \
      m->field = v;                                                             \
    });                                                                         \
  }

HH_SETTER(epsilon, double, epsilon, if (!(v >= 0)) throw HHError(HH_ERR_ARG, "epsilon < 0"))
HH_SETTER(eta, double, eta, (void)0)
HH_SETTER(max_iter, int, max_iter,
          if (v < 1 || v > 128) throw HHError(HH_ERR_ARG, "max_iter must be in [1, 128]"))
HH_SETTER(mat_pool_size_factor, int, pool_factor, (void)0)
HH_SETTER(max_far_dot_work_node_size, int, far_work, (void)0)
HH_SETTER(max_far_field_distance, double, max_far_dist, (void)0)

int hh_model_set_max_relative_coord(hh_model *m, int64_t v) {
  return guarded([&]() {
    if (!m) throw HHError(HH_ERR_ARG, "NULL model");
    if (v < 1) throw HHError(HH_ERR_ARG, "max_relative_coord must be positive");
    m->max_rel_coord = v;
  });
}

int hh_model_set_post_check_flag(hh_model *m, int v) {
  return guarded([&]() {
    if (!m) throw HHError(HH_ERR_ARG, "NULL model");
    m->post_check = v != 0;
  });
}

int hh_model_set_use_symmetry_flag(hh_model *m, int v) {
  return guarded([&]() {
    if (!m) throw HHError(HH_ERR_ARG, "NULL model");
    if (m->use_symmetry != (v != 0)) m->built = false;  // lists depend on it
    m->use_symmetry = v != 0;
  });
}

int hh_model_set_precision(hh_model *m, int p) {
  return guarded([&]() {
    if (!m) throw HHError(HH_ERR_ARG, "NULL model");
    if (p != HH_PREC_FP32 && p != HH_PREC_BF16 && p != HH_PREC_FP16)
      throw HHError(HH_ERR_ARG, "unknown precision");
    m->prec = p;
  });
}

int hh_model_set_tree_reuse(hh_model *m, int flag) {
  return guarded([&]() {
    if (!m) throw HHError(HH_ERR_ARG, "NULL model");
    m->tree_reuse = flag != 0;
  });
}

int hh_model_update_coord(hh_model *m, const float *coord, int64_t n) {
  return guarded([&]() {
    if (!m || !coord) throw HHError(HH_ERR_ARG, "NULL argument");
    if (n != m->n) throw HHError(HH_ERR_ARG, "particle count differs from the model's");
    m->x.assign(coord, coord + 3 * n);
    m->x_stale = false;
    m->update();
  });
}

int hh_model_update_coord_device(hh_model *m, const float *coord_dev, int64_t n, void *stream) {
  return guarded([&]() {
    if (!m || !coord_dev) throw HHError(HH_ERR_ARG, "NULL argument");
    if (n != m->n) throw HHError(HH_ERR_ARG, "particle count differs from the model's");
    const hipStream_t st = stream ? (hipStream_t)stream : m->stream;
    if (m->tree_reuse && m->built) {
      // kept tree: positions in tree order straight from the device array
      HIPCHK(launch_gather_forces(coord_dev, n, m->perm.as<int32_t>(), m->xp.as<float4>(), st));
      m->x_stale = true;  // the host copy catches up before the next rebuild
      return;
    }
    if (m->tree_device) {
      // build from the device coordinates; the host copy is refreshed from
      // xp when a host path needs it
      HIPCHK(hipStreamSynchronize(st));
      m->update(coord_dev);
      return;
    }
    m->x.resize(3 * (size_t)n);
    HIPCHK(hipMemcpyAsync(m->x.data(), coord_dev, 12 * (size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    m->x_stale = false;
    m->update();
  });
}

int hh_model_set_near_overlap(hh_model *m, int on) {
  return guarded([&]() {
    if (!m) throw HHError(HH_ERR_ARG, "NULL model");
    if (on != 0 && on != 1) throw HHError(HH_ERR_ARG, "near overlap: 0 or 1");
    m->near_overlap = on != 0;
  });
}

int hh_model_set_tree_build(hh_model *m, int device) {
  return guarded([&]() {
    if (!m) throw HHError(HH_ERR_ARG, "NULL model");
    if (device != 0 && device != 1) throw HHError(HH_ERR_ARG, "tree build: 0 host, 1 device");
    m->tree_device = device != 0;
  });
}

int hh_model_tune_far_field(hh_model *m, const float *f_dev, int64_t n, void *stream,
                            double *table, int max_rows, int *rows) {
  return guarded([&]() {
    if (!m || !f_dev || !rows) throw HHError(HH_ERR_ARG, "NULL argument");
    if (n != m->n) throw HHError(HH_ERR_ARG, "array length differs from the model's");
    std::vector<std::array<double, 3>> out;
    m->tune_far_field(f_dev, stream ? (hipStream_t)stream : m->stream, out);
    const int nr = (int)std::min<size_t>(out.size(), (size_t)std::max(max_rows, 0));
    for (int r = 0; r < nr && table; r++)
      for (int c = 0; c < 3; c++) table[3 * r + c] = out[(size_t)r][(size_t)c];
    *rows = nr;
  });
}

int hh_model_set_near_shape(hh_model *m, int shape) {
  return guarded([&]() {
    if (!m) throw HHError(HH_ERR_ARG, "NULL model");
    if (shape < 0 || shape >= kNearShapeCount) throw HHError(HH_ERR_ARG, "near shape out of range");
    m->near_hcfg = shape;
  });
}

int hh_model_set_near_workgroups(hh_model *m, int workgroups) {
  return guarded([&]() {
    if (!m) throw HHError(HH_ERR_ARG, "NULL model");
    if (workgroups < 0) throw HHError(HH_ERR_ARG, "negative workgroup count");
    m->near_cus = workgroups;
  });
}

int hh_model_tune_near(hh_model *m, const float *f_dev, int64_t n, void *stream, double *table,
                       int max_rows, int *rows) {
  return guarded([&]() {
    if (!m || !f_dev || !rows) throw HHError(HH_ERR_ARG, "NULL argument");
    if (n != m->n) throw HHError(HH_ERR_ARG, "array length differs from the model's");
    std::vector<std::array<double, 5>> out;
    m->tune_near(f_dev, stream ? (hipStream_t)stream : m->stream, out);
    const int nr = (int)std::min<size_t>(out.size(), (size_t)std::max(max_rows, 0));
    for (int r = 0; r < nr && table; r++)
      for (int c = 0; c < 5; c++) table[5 * r + c] = out[(size_t)r][(size_t)c];
    *rows = nr;
  });
}

int hh_model_dot_capturable(hh_model *m, int *capturable, uint64_t *signature) {
  return guarded([&]() {
    if (!m || !capturable || !signature) throw HHError(HH_ERR_ARG, "NULL argument");
    *capturable = m->capturable() ? 1 : 0;
    *signature = m->graph_signature();
  });
}

int hh_model_dot_enqueue(hh_model *m, float *u_dev, const float *f_dev, int64_t n, void *stream) {
  return guarded([&]() {
    if (!m || !u_dev || !f_dev) throw HHError(HH_ERR_ARG, "NULL argument");
    if (n != m->n) throw HHError(HH_ERR_ARG, "array length differs from the model's");
    m->enqueue_dot(u_dev, f_dev, stream ? (hipStream_t)stream : m->stream);
  });
}

int hh_model_dot(hh_model *m, float *u, const float *f, int64_t n) {
  return guarded([&]() {
    if (!m || !u || !f) throw HHError(HH_ERR_ARG, "NULL argument");
    if (n != m->n) throw HHError(HH_ERR_ARG, "array length differs from the model's");
    m->dot_host(u, f, false);
  });
}

int hh_model_dense_dot(hh_model *m, float *u, const float *f, int64_t n) {
  return guarded([&]() {
    if (!m || !u || !f) throw HHError(HH_ERR_ARG, "NULL argument");
    if (n != m->n) throw HHError(HH_ERR_ARG, "array length differs from the model's");
    m->dot_host(u, f, true);
  });
}

int hh_model_dot_device(hh_model *m, float *u_dev, const float *f_dev, int64_t n, void *stream) {
  return guarded([&]() {
    if (!m || !u_dev || !f_dev) throw HHError(HH_ERR_ARG, "NULL argument");
    if (n != m->n) throw HHError(HH_ERR_ARG, "array length differs from the model's");
    m->dot_device(u_dev, f_dev, stream ? (hipStream_t)stream : m->stream);
  });
}

int hh_model_tree_counts(hh_model *m, int64_t counts[4]) {
  return guarded([&]() {
    if (!m || !counts) throw HHError(HH_ERR_ARG, "NULL argument");
    if (!m->built) throw HHError(HH_ERR_STATE, "update_coord has not run");
    counts[0] = m->tl.T;
    counts[1] = (int64_t)m->tl.closeI.size();
    counts[2] = (int64_t)m->tl.farI.size();
    counts[3] = (int64_t)m->tl.leaves.size();
  });
}

int hh_model_tree_export(hh_model *m, uint64_t *tree5, uint64_t *perm, float *aux6,
                         int64_t *close_ij, int64_t *far_ij, int64_t *leaves) {
  return guarded([&]() {
    if (!m) throw HHError(HH_ERR_ARG, "NULL model");
    if (!m->built) throw HHError(HH_ERR_STATE, "update_coord has not run");
    export_lists(m->tl, tree5, perm, aux6, close_ij, far_ij, leaves);
  });
}

int hh_model_far_ranks(hh_model *m, int32_t *ranks) {
  return guarded([&]() {
    if (!m || !ranks) throw HHError(HH_ERR_ARG, "NULL argument");
    m->fetch_ranks();
    std::copy(m->far_rank.begin(), m->far_rank.end(), ranks);
  });
}

int hh_model_check_fp64(hh_model *m, const float *f, const int64_t *targets, int64_t nt,
                        int convention, double *u_out) {
  return guarded([&]() {
    if (!m || !f || !targets || !u_out) throw HHError(HH_ERR_ARG, "NULL argument");
    if (nt <= 0) throw HHError(HH_ERR_ARG, "need at least one target");
    if (convention != 0 && convention != 1) throw HHError(HH_ERR_ARG, "convention must be 0 or 1");
    m->check_fp64(f, targets, nt, convention, u_out);
  });
}

int hh_model_eval_pair(hh_model *m, const float *r, int64_t P, float *out9, int prec) {
  return guarded([&]() {
    if (!m || !r || !out9) throw HHError(HH_ERR_ARG, "NULL argument");
    if (P <= 0) throw HHError(HH_ERR_ARG, "need P > 0");
    if (prec < 0 || prec > 3) throw HHError(HH_ERR_ARG, "unknown precision");
    if (!m->has_weights) throw HHError(HH_ERR_STATE, "two-body model not loaded");
    DBuf dr, dout;
    dr.reserve(12 * (size_t)P);
    dout.reserve(36 * (size_t)P);
    try {
      HIPCHK(hipMemcpy(dr.p, r, 12 * (size_t)P, hipMemcpyHostToDevice));
      HIPCHK(launch_pair(prec, dr.as<float>(), P, dout.as<float>(), m->devw_f(prec == PREC_FP64),
                         m->devw_h(prec), g_cus, m->stream));
      HIPCHK(hipStreamSynchronize(m->stream));
      HIPCHK(hipMemcpy(out9, dout.p, 36 * (size_t)P, hipMemcpyDeviceToHost));
    } catch (...) {
      dr.release();
      dout.release();
      throw;
    }
    dr.release();
    dout.release();
  });
}

int hh_model_surrogate_samples(hh_model *m, int64_t n, float r_lo, float r_hi, uint64_t seed,
                               float *r_dev, float *t_dev, void *stream) {
  return guarded([&]() {
    if (!m || !r_dev || !t_dev) throw HHError(HH_ERR_ARG, "NULL argument");
    if (n <= 0) throw HHError(HH_ERR_ARG, "need n > 0");
    if (!(r_lo > 0) || !(r_hi > r_lo)) throw HHError(HH_ERR_ARG, "need 0 < r_lo < r_hi");
    if (!m->has_weights) throw HHError(HH_ERR_STATE, "two-body model not loaded");
    hipStream_t st = stream ? (hipStream_t)stream : m->stream;
    HIPCHK(launch_sur_samples(n, r_lo, r_hi, seed, r_dev, st));
    // label in chunks through a bounded block buffer
    const int64_t chunk = std::min<int64_t>(n, (int64_t)1 << 20);
    DBuf m9;
    m9.reserve(36 * (size_t)chunk);
    try {
      for (int64_t k0 = 0; k0 < n; k0 += chunk) {
        const int64_t c = std::min(chunk, n - k0);
        HIPCHK(launch_pair(PREC_FP32, r_dev + 3 * k0, c, m9.as<float>(), m->devw_f(false),
                           m->devw_h(PREC_BF16), g_cus, st));
        HIPCHK(launch_sur_targets(r_dev + 3 * k0, m9.as<float>(), c, t_dev + 6 * k0, st));
      }
      HIPCHK(hipStreamSynchronize(st));
    } catch (...) {
      m9.release();
      throw;
    }
    m9.release();
  });
}

int hh_model_set_surrogate(hh_model *m, hh_surrogate *s, double tol, int samples) {
  return guarded([&]() {
    if (!m) throw HHError(HH_ERR_ARG, "NULL model");
    if (s && !(tol >= 0)) throw HHError(HH_ERR_ARG, "tol must be >= 0");
    if (s && (samples < 1 || samples > 4096)) throw HHError(HH_ERR_ARG, "samples must be in 1..4096");
    m->sur = s;
    if (s) m->sur_tol = tol, m->sur_samples = samples;
  });
}

int hh_model_surrogate_stats(hh_model *m, int64_t stats[5]) {
  return guarded([&]() {
    if (!m || !stats) throw HHError(HH_ERR_ARG, "NULL argument");
    m->fetch_judge();
    for (int k = 0; k < 5; k++) stats[k] = m->sur_stats[k];
  });
}

int hh_model_surrogate_errors(hh_model *m, float *err) {
  return guarded([&]() {
    if (!m || !err) throw HHError(HH_ERR_ARG, "NULL argument");
    m->fetch_judge();
    if (m->sur_err.size() != m->far.size())
      throw HHError(HH_ERR_STATE, "no judged matvec since the last tree build");
    std::copy(m->sur_err.begin(), m->sur_err.end(), err);
  });
}

struct hh_tree {
  TreeLists tl;
};

int hh_tree_build(const float *x, int64_t n, int block_size, int use_symmetry,
                  int64_t max_relative_coord, hh_tree **out) {
  return guarded([&]() {
    if (!x || !out) throw HHError(HH_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (n <= 0 || block_size <= 0 || max_relative_coord <= 0)
      throw HHError(HH_ERR_ARG, "need n, block_size, max_relative_coord > 0");
    hh_tree *t = new hh_tree();
    build_tree_lists(x, n, block_size, use_symmetry != 0, max_relative_coord, 0, t->tl);
    *out = t;
  });
}

int hh_tree_counts(hh_tree *t, int64_t counts[4]) {
  return guarded([&]() {
    if (!t || !counts) throw HHError(HH_ERR_ARG, "NULL argument");
    counts[0] = t->tl.T;
    counts[1] = (int64_t)t->tl.closeI.size();
    counts[2] = (int64_t)t->tl.farI.size();
    counts[3] = (int64_t)t->tl.leaves.size();
  });
}

static void export_lists(const TreeLists &t, uint64_t *tree5, uint64_t *perm, float *aux6,
                         int64_t *close_ij, int64_t *far_ij, int64_t *leaves) {
  if (tree5) std::memcpy(tree5, t.rows.data(), t.rows.size() * 8);
  if (perm)
    for (int64_t i = 0; i < t.n; i++) perm[i] = (uint64_t)t.perm[(size_t)i];
  if (aux6) std::memcpy(aux6, t.aux.data(), t.aux.size() * 4);
  if (close_ij)
    for (size_t b = 0; b < t.closeI.size(); b++) close_ij[2 * b] = t.closeI[b], close_ij[2 * b + 1] = t.closeJ[b];
  if (far_ij)
    for (size_t b = 0; b < t.farI.size(); b++) far_ij[2 * b] = t.farI[b], far_ij[2 * b + 1] = t.farJ[b];
  if (leaves) std::memcpy(leaves, t.leaves.data(), t.leaves.size() * 8);
}

int hh_tree_export(hh_tree *t, uint64_t *tree5, uint64_t *perm, float *aux6, int64_t *close_ij,
                   int64_t *far_ij, int64_t *leaves) {
  return guarded([&]() {
    if (!t) throw HHError(HH_ERR_ARG, "NULL tree");
    export_lists(t->tl, tree5, perm, aux6, close_ij, far_ij, leaves);
  });
}

int hh_tree_build_device(const float *x_dev, const int64_t *counts, int64_t ninst, int block_size,
                         int use_symmetry, int64_t max_relative_coord, void *stream, hh_tree **out) {
  return guarded([&]() {
    if (!x_dev || !counts || !out) throw HHError(HH_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (ninst <= 0 || block_size <= 0 || max_relative_coord <= 0)
      throw HHError(HH_ERR_ARG, "need ninst, block_size, max_relative_coord > 0");
    std::vector<int64_t> pb{0};
    for (int64_t k = 0; k < ninst; k++) {
      if (counts[k] <= 0) throw HHError(HH_ERR_ARG, "instance sizes must be > 0");
      pb.push_back(pb.back() + counts[k]);
    }
    if (pb.back() >= ((int64_t)1 << 31)) throw HHError(HH_ERR_ARG, "too many particles");
    ensure_device();
    std::unique_ptr<hh_tree> t(new hh_tree());
    DeviceTreeBuilder *b = device_tree_builder_new();
    DBuf permd;
    try {
      permd.reserve(4 * (size_t)pb.back());
      build_tree_lists_device(b, x_dev, pb, block_size, use_symmetry != 0, max_relative_coord,
                              stream, t->tl, permd.as<int32_t>());
    } catch (...) {
      permd.release();
      device_tree_builder_free(b);
      throw;
    }
    permd.release();
    device_tree_builder_free(b);
    *out = t.release();
  });
}

void hh_tree_free(hh_tree *t) { delete t; }

int hh_model_last_profile(hh_model *m, double ms[2], int64_t pairs[2], int64_t launches[2]) {
  return guarded([&]() {
    if (!m || !ms || !pairs || !launches) throw HHError(HH_ERR_ARG, "null argument");
    for (int i = 0; i < 2; i++) {
      ms[i] = m->prof_ms[i];
      pairs[i] = m->prof_pairs[i];
      launches[i] = m->prof_launches[i];
    }
  });
}

int hh_model_last_timings(hh_model *m, double t[3]) {
  return guarded([&]() {
    if (!m || !t) throw HHError(HH_ERR_ARG, "NULL argument");
    t[0] = m->t_close;
    t[1] = m->t_far;
    t[2] = m->t_total;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Device memory (include/hh_device.h)
// ---------------------------------------------------------------------------
extern "C" {

int hh_device_alloc(size_t bytes, void **p) {
  return guarded([&]() {
    if (!p) throw HHError(HH_ERR_ARG, "NULL argument");
    *p = nullptr;
    ensure_device();
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
      *p = nullptr;
      throw HHError(HH_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
  });
}

int hh_device_free(void *p) {
  return guarded([&]() {
    if (p) HIPCHK(hipFree(p));
  });
}

int hh_memcpy_to_device(void *dst_dev, const void *src_host, size_t bytes) {
  return guarded([&]() {
    if ((!dst_dev || !src_host) && bytes) throw HHError(HH_ERR_ARG, "NULL argument");
    if (bytes) HIPCHK(hipMemcpy(dst_dev, src_host, bytes, hipMemcpyHostToDevice));
  });
}

int hh_memcpy_to_host(void *dst_host, const void *src_dev, size_t bytes) {
  return guarded([&]() {
    if ((!dst_host || !src_dev) && bytes) throw HHError(HH_ERR_ARG, "NULL argument");
    if (bytes) HIPCHK(hipMemcpy(dst_host, src_dev, bytes, hipMemcpyDeviceToHost));
  });
}

int hh_device_copy(void *dst_dev, const void *src_dev, size_t bytes) {
  return guarded([&]() {
    if ((!dst_dev || !src_dev) && bytes) throw HHError(HH_ERR_ARG, "NULL argument");
    if (bytes % 16) throw HHError(HH_ERR_ARG, "bytes must be a multiple of 16");
    ensure_device();
    HIPCHK(launch_copy16(dst_dev, src_dev, bytes / 16, nullptr));
    HIPCHK(hipDeviceSynchronize());
  });
}

int hh_device_synchronize(void) {
  return guarded([&]() {
    ensure_device();
    HIPCHK(hipDeviceSynchronize());
  });
}

}  // extern "C"
