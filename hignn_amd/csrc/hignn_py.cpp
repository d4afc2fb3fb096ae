// pybind11 module `hignn`: the reference's Python surface
// (interface/HignnInterface.cpp:9-92) on top of the C ABI of libhh.
//
// HignnModel keeps the reference's method names, argument order and types.
// dot/dense_dot write into the caller's float32 (N,3) array in place
// (src/HignnModel.cpp:570, 618-624); a wrongly typed or shaped output array
// raises instead of being silently copied.  ExplicitEuler and ExplicitRk4
// restate include/TimeIntegrator.hpp:431-860 (see the notes on each).

#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hh.h"
#include "../../include/hh_dynamics.h"
#include "../../include/hh_three_body.h"

namespace py = pybind11;

namespace {

void check(int rc) {
  if (rc != HH_OK) throw std::runtime_error(std::string("hignn: ") + hh_last_error());
}

// (N,3) float32 copy of any array-like (honours strides, e.g. x[:, 0:3]).
std::vector<float> rows3(const py::array_t<float, py::array::forcecast> &a, const char *what) {
  if (a.ndim() != 2 || a.shape(1) < 3)
    throw std::invalid_argument(std::string(what) + " must be a 2-D array with 3 columns");
  auto v = a.unchecked<2>();
  std::vector<float> out(3 * (size_t)a.shape(0));
  for (py::ssize_t i = 0; i < a.shape(0); i++)
    for (int d = 0; d < 3; d++) out[3 * (size_t)i + d] = v(i, d);
  return out;
}

void write_rows3(py::array &u, const std::vector<float> &src, py::ssize_t n) {
  if (!py::isinstance<py::array_t<float>>(u) || !u.dtype().is(py::dtype::of<float>()))
    throw py::type_error("u must be a float32 numpy array (it is written in place)");
  if (u.ndim() != 2 || u.shape(0) != n || u.shape(1) < 3)
    throw std::invalid_argument("u must have shape (N, 3)");
  if (!u.writeable()) throw std::invalid_argument("u is read-only");
  auto a = py::array_t<float>(u);
  auto v = a.mutable_unchecked<2>();
  for (py::ssize_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) v(i, d) = src[3 * (size_t)i + d];
}

class Model {
 public:
  Model(py::array_t<float, py::array::forcecast> coord, int block_size) {
    auto x = rows3(coord, "coord");
    n_ = coord.shape(0);
    check(hh_model_create(x.data(), n_, block_size, &m_));
  }
  ~Model() { hh_model_destroy(m_); }
  hh_model *handle() const { return m_; }
  int64_t size() const { return n_; }

  // load_two_body_model(base): base_<device>.pt or base.pt (TorchScript, as
  // src/HignnModel.cpp:404-411), base.pkl (torch-saved Sequential), base.npy,
  // read without unpickling; or "xavier:<seed>" / "smooth:<seed>" for the
  // seeded weight sets of the same layer shapes (hignn_amd/weights.py).
  void load_two_body_model(const std::string &path) {
    py::object W = py::module_::import("hignn_amd.weights");
    py::array_t<float, py::array::forcecast | py::array::c_style> w;
    if (path.rfind("xavier:", 0) == 0)
      w = W.attr("xavier")(std::stoi(path.substr(7)));
    else
      w = W.attr("load")(path);
    load_two_body_weights(w);
  }
  void load_two_body_weights(py::array_t<float, py::array::forcecast | py::array::c_style> w) {
    check(hh_model_load_two_body(m_, w.data(), (int64_t)w.size()));
  }
  void load_three_body_model(const std::string &) {}  // the reference's is a no-op too

  void set_epsilon(double v) { check(hh_model_set_epsilon(m_, v)); }
  void set_eta(double v) { check(hh_model_set_eta(m_, v)); }
  void set_max_iter(int v) { check(hh_model_set_max_iter(m_, v)); }
  void set_mat_pool_size_factor(int v) { check(hh_model_set_mat_pool_size_factor(m_, v)); }
  void set_max_far_dot_work_node_size(int v) { check(hh_model_set_max_far_dot_work_node_size(m_, v)); }
  void set_max_relative_coord(size_t v) { check(hh_model_set_max_relative_coord(m_, (int64_t)v)); }
  void set_post_check_flag(bool v) { check(hh_model_set_post_check_flag(m_, v)); }
  void set_tree_reuse(bool v) { check(hh_model_set_tree_reuse(m_, v)); }
  void set_tree_build(const std::string &where) {
    if (where != "device" && where != "host") throw std::invalid_argument("tree build: 'device' or 'host'");
    check(hh_model_set_tree_build(m_, where == "device" ? 1 : 0));
  }
  void set_use_symmetry_flag(bool v) { check(hh_model_set_use_symmetry_flag(m_, v)); }
  void set_max_far_field_distance(double v) { check(hh_model_set_max_far_field_distance(m_, v)); }

  void set_precision(const std::string &p) {
    int v;
    if (p == "fp32") v = HH_PREC_FP32;
    else if (p == "bf16") v = HH_PREC_BF16;
    else if (p == "fp16") v = HH_PREC_FP16;
    else throw std::invalid_argument("precision must be fp32, bf16 or fp16");
    check(hh_model_set_precision(m_, v));
  }

  void update_coord(py::array_t<float, py::array::forcecast> coord) {
    auto x = rows3(coord, "coord");
    if (coord.shape(0) != n_) throw std::invalid_argument("coordinate count differs from the model's");
    check(hh_model_update_coord(m_, x.data(), n_));
  }

  void dot(py::array &u, py::array_t<float, py::array::forcecast> f) { apply(u, f, false); }
  void dense_dot(py::array &u, py::array_t<float, py::array::forcecast> f) { apply(u, f, true); }

  // u, f are device pointers (e.g. torch tensor data_ptr()); stream 0 = own.
  void dot_device(uintptr_t u, uintptr_t f, uintptr_t stream) {
    py::gil_scoped_release nogil;
    check(hh_model_dot_device(m_, (float *)u, (const float *)f, n_, (void *)stream));
  }

  // Far-field launch tuning on device forces f (a data pointer); returns the
  // rows (mode, chunk, far-field ms) and keeps the fastest candidate.
  py::list tune_far_field(uintptr_t f, uintptr_t stream) {
    double t[8 * 3];
    int rows = 0;
    {
      py::gil_scoped_release nogil;
      check(hh_model_tune_far_field(m_, (const float *)f, n_, (void *)stream, t, 8, &rows));
    }
    py::list out;
    for (int r = 0; r < rows; r++)
      out.append(py::make_tuple((int)t[3 * r], (int)t[3 * r + 1], t[3 * r + 2]));
    return out;
  }

  // Launch tuning of the 16-bit near kernel on device forces f; returns the
  // rows (shape, tiles, threads, workgroups, matvec ms) and keeps the fastest.
  py::list tune_near(uintptr_t f, uintptr_t stream) {
    double t[16 * 5];
    int rows = 0;
    {
      py::gil_scoped_release nogil;
      check(hh_model_tune_near(m_, (const float *)f, n_, (void *)stream, t, 16, &rows));
    }
    py::list out;
    for (int r = 0; r < rows; r++)
      out.append(py::make_tuple((int)t[5 * r], (int)t[5 * r + 1], (int)t[5 * r + 2],
                                (int)t[5 * r + 3], t[5 * r + 4]));
    return out;
  }
  void set_near_shape(int shape) { check(hh_model_set_near_shape(m_, shape)); }
  void set_near_workgroups(int wg) { check(hh_model_set_near_workgroups(m_, wg)); }

  py::dict tree_export() {
    int64_t c[4];
    check(hh_model_tree_counts(m_, c));
    py::array_t<uint64_t> tree({(py::ssize_t)c[0], (py::ssize_t)5});
    py::array_t<uint64_t> perm(n_);
    py::array_t<float> aux({(py::ssize_t)c[0], (py::ssize_t)6});
    py::array_t<int64_t> close({(py::ssize_t)c[1], (py::ssize_t)2});
    py::array_t<int64_t> far({(py::ssize_t)c[2], (py::ssize_t)2});
    py::array_t<int64_t> leaves(c[3]);
    check(hh_model_tree_export(m_, tree.mutable_data(), perm.mutable_data(), aux.mutable_data(),
                               close.mutable_data(), far.mutable_data(), leaves.mutable_data()));
    py::dict d;
    d["tree"] = tree;
    d["perm"] = perm;
    d["aux"] = aux;
    d["close"] = close;
    d["far"] = far;
    d["leaves"] = leaves;
    return d;
  }

  py::array_t<int32_t> far_ranks() {
    int64_t c[4];
    check(hh_model_tree_counts(m_, c));
    py::array_t<int32_t> r(c[2]);
    check(hh_model_far_ranks(m_, r.mutable_data()));
    return r;
  }

  py::array_t<double> check_fp64(py::array_t<float, py::array::forcecast> f,
                                 py::array_t<int64_t, py::array::forcecast> targets, int convention) {
    auto fv = rows3(f, "f");
    py::array_t<double> out({(py::ssize_t)targets.size(), (py::ssize_t)3});
    check(hh_model_check_fp64(m_, fv.data(), targets.data(), (int64_t)targets.size(), convention,
                              out.mutable_data()));
    return out;
  }

  py::tuple last_timings() {
    double t[3];
    check(hh_model_last_timings(m_, t));
    return py::make_tuple(t[0], t[1], t[2]);
  }

  py::dict last_profile() {
    double ms[2];
    int64_t pairs[2], launches[2];
    check(hh_model_last_profile(m_, ms, pairs, launches));
    py::dict d;
    d["near"] = py::make_tuple(ms[0], pairs[0], launches[0]);
    d["far_mlp"] = py::make_tuple(ms[1], pairs[1], launches[1]);
    return d;
  }

 private:
  void apply(py::array &u, py::array_t<float, py::array::forcecast> f, bool dense) {
    auto fv = rows3(f, "f");
    if (f.shape(0) != n_) throw std::invalid_argument("f has the wrong number of rows");
    std::vector<float> uv(3 * (size_t)n_);
    {
      py::gil_scoped_release nogil;
      check(dense ? hh_model_dense_dot(m_, uv.data(), fv.data(), n_)
                  : hh_model_dot(m_, uv.data(), fv.data(), n_));
    }
    write_rows3(u, uv, n_);
  }

  hh_model *m_ = nullptr;
  py::ssize_t n_ = 0;
};

// ---------------------------------------------------------------------------
// Forcing on the device (include/hh_dynamics.h).  PotentialForce keeps the
// reference's class (interface/HignnInterface.cpp:66-71, PotentialForce.hpp):
// get_potential_force(coord) returns the Morse force of every particle.
// DeviceForcing adds the body force and filament terms of
// python/simulate.py:90-144 and is what set_device_velocity runs per stage.
// ---------------------------------------------------------------------------
class Forcing {
 public:
  Forcing() { check(hh_forcing_create(&f_)); }
  ~Forcing() { hh_forcing_destroy(f_); }
  hh_forcing *handle() const { return f_; }
  void set_periodic(bool flag) { check(hh_forcing_set_periodic(f_, flag)); }
  void set_domain(py::array_t<float, py::array::forcecast> domain) {
    if (domain.ndim() != 2) throw std::runtime_error("Rank of domain must be two");
    if (domain.shape(0) < 2 || domain.shape(1) < 3)
      throw std::invalid_argument("domain must have shape (2, 3)");
    auto d = domain.unchecked<2>();
    const float lo[3] = {d(0, 0), d(0, 1), d(0, 2)}, hi[3] = {d(1, 0), d(1, 1), d(1, 2)};
    check(hh_forcing_set_domain(f_, lo, hi));
  }
  void set_two_body_epsilon(float eps) { check(hh_forcing_set_morse(f_, eps)); }
  void set_body_force(std::vector<float> b) {
    if (b.size() != 3) throw std::invalid_argument("body force must have 3 components");
    check(hh_forcing_set_body(f_, b.data()));
  }
  void set_filaments(int64_t n_filament, int64_t n_chain, float rest_length, float k_t,
                     double k_b) {
    check(hh_forcing_set_filaments(f_, n_filament, n_chain, rest_length, k_t, k_b));
  }
  void set_bonds(int64_t n_filament, int64_t n_chain, double D, double alpha, double r0) {
    check(hh_forcing_set_bonds(f_, n_filament, n_chain, D, alpha, r0));
  }
  py::array_t<float> eval(py::array_t<float, py::array::forcecast> coord) {
    if (coord.ndim() != 2) throw std::runtime_error("Rank of target sites must be two");
    auto x = rows3(coord, "coord");
    const py::ssize_t n = coord.shape(0);
    py::array_t<float> out({n, (py::ssize_t)3});
    check(hh_forcing_eval(f_, x.data(), out.mutable_data(), (int64_t)n));
    return out;
  }

 private:
  hh_forcing *f_ = nullptr;
};

class PotentialForceCls : public Forcing {};

// ---------------------------------------------------------------------------
// NeighborLists (interface/HignnInterface.cpp:73-91, include/NeighborLists.hpp)
// on the device lists of include/hh_three_body.h.  Edge arrays are uint64
// (the reference's size_t), (2,E) or (3,E).  Inside one target the edges come
// ascending by neighbour; the reference's order is its kd-tree's.
// ---------------------------------------------------------------------------
class Neighbors {
 public:
  Neighbors() { check(hh_neighbors_create(&nb_)); }
  ~Neighbors() { hh_neighbors_destroy(nb_); }
  void set_periodic(bool flag) { check(hh_neighbors_set_periodic(nb_, flag)); }
  void set_domain(py::array_t<float, py::array::forcecast> domain) {
    if (domain.ndim() != 2) throw std::runtime_error("Rank of domain must be two");
    if (domain.shape(0) < 2 || domain.shape(1) < 3)
      throw std::invalid_argument("domain must have shape (2, 3)");
    auto d = domain.unchecked<2>();
    const float lo[3] = {d(0, 0), d(0, 1), d(0, 2)}, hi[3] = {d(1, 0), d(1, 1), d(1, 2)};
    check(hh_neighbors_set_domain(nb_, lo, hi));
  }
  void update_coord(py::array_t<float, py::array::forcecast> coord) {
    if (coord.ndim() != 2) throw std::runtime_error("Rank of target sites must be two");
    auto x = rows3(coord, "coord");
    n_ = coord.shape(0);
    check(hh_neighbors_update_coord(nb_, x.data(), n_));
  }
  void set_two_body_epsilon(float e) { check(hh_neighbors_set_two_body_epsilon(nb_, e)); }
  void set_three_body_epsilon(float e) { check(hh_neighbors_set_three_body_epsilon(nb_, e)); }
  void build_two_body_info() { check(hh_neighbors_build_two_body(nb_, &e2_)); }
  void build_three_body_info() { check(hh_neighbors_build_three_body(nb_, &e3_, &es_)); }

  py::array_t<uint64_t> two_body() { return fetch(2, e2_, hh_neighbors_two_body_edges); }
  py::array_t<uint64_t> three_body() { return fetch(3, e3_, hh_neighbors_three_body_edges); }
  py::array_t<uint64_t> three_body_self() {
    return fetch(2, es_, hh_neighbors_three_body_self_edges);
  }

  // get_two_body_edge_info_by_index (NeighborLists.hpp:309-340): every pair
  // (j, index), j != index, over all particles.
  py::array_t<uint64_t> two_body_by_index(py::array_t<uint64_t, py::array::forcecast> idx) {
    auto ix = idx.unchecked<1>();
    const int64_t nt = idx.shape(0), per = n_ > 0 ? n_ - 1 : 0, ne = nt * per;
    py::array_t<uint64_t> out({(py::ssize_t)2, (py::ssize_t)ne});
    uint64_t *o = out.mutable_data();
    for (int64_t i = 0; i < nt; i++) {
      int64_t c = i * per;
      for (int64_t j = 0; j < n_; j++)
        if ((uint64_t)j != ix(i)) {
          o[c] = (uint64_t)j;
          o[ne + c] = ix(i);
          c++;
        }
    }
    return out;
  }

  // get_three_body_edge_info_by_index (:362-403): the edges of each given
  // target, rows (p, n, i); get_three_body_edge_self_info_by_index
  // (:424-464): rows (n, i).
  py::array_t<uint64_t> three_body_by_index(py::array_t<uint64_t, py::array::forcecast> idx) {
    return by_index(idx, three_body(), 3);
  }
  py::array_t<uint64_t> three_body_self_by_index(py::array_t<uint64_t, py::array::forcecast> idx) {
    return by_index(idx, three_body_self(), 2);
  }

 private:
  py::array_t<uint64_t> fetch(int rows, int64_t ne, int (*fn)(hh_neighbors *, int64_t *)) {
    std::vector<int64_t> buf((size_t)rows * (size_t)ne);
    if (ne > 0) check(fn(nb_, buf.data()));
    py::array_t<uint64_t> out({(py::ssize_t)rows, (py::ssize_t)ne});
    uint64_t *o = out.mutable_data();
    for (size_t k = 0; k < buf.size(); k++) o[k] = (uint64_t)buf[k];
    return out;
  }
  // edges of a target are contiguous (row 0 ascending); rows are reversed
  py::array_t<uint64_t> by_index(py::array_t<uint64_t, py::array::forcecast> idx,
                                 py::array_t<uint64_t> all, int rows) {
    auto ix = idx.unchecked<1>();
    const int64_t ne = all.shape(1);
    const uint64_t *a = all.data();
    std::vector<std::pair<int64_t, int64_t>> span;
    int64_t tot = 0;
    for (py::ssize_t i = 0; i < idx.shape(0); i++) {
      const int64_t lo = std::lower_bound(a, a + ne, ix(i)) - a;
      const int64_t hi = std::upper_bound(a, a + ne, ix(i)) - a;
      span.emplace_back(lo, hi);
      tot += hi - lo;
    }
    py::array_t<uint64_t> out({(py::ssize_t)rows, (py::ssize_t)tot});
    uint64_t *o = out.mutable_data();
    int64_t c = 0;
    for (auto &sp : span)
      for (int64_t e = sp.first; e < sp.second; e++, c++)
        for (int r = 0; r < rows; r++) o[r * tot + c] = a[(rows - 1 - r) * ne + e];
    return out;
  }

  hh_neighbors *nb_ = nullptr;
  int64_t n_ = 0, e2_ = 0, e3_ = 0, es_ = 0;
};

// Three-body and self corrections (velocity_update_without_2body,
// python/gravity_field.py:34-59) on the device.
class ThreeBody {
 public:
  ThreeBody() { check(hh_three_body_create(&tb_)); }
  ~ThreeBody() { hh_three_body_destroy(tb_); }
  hh_three_body *handle() const { return tb_; }
  // paths of the torch-saved nets (HIGNN_nn_3body.pkl, HIGNN_nn_self.pkl),
  // read without unpickling, .npy vectors or "xavier:<seed>"
  void load_model(const std::string &three_body, const std::string &self_net) {
    py::object W = py::module_::import("hignn_amd.weights");
    py::array_t<float, py::array::forcecast | py::array::c_style> a =
        W.attr("load_net")(three_body, 6);
    py::array_t<float, py::array::forcecast | py::array::c_style> b =
        W.attr("load_net")(self_net, 3);
    check(hh_three_body_load(tb_, a.data(), a.size(), b.data(), b.size()));
  }
  void load_weights(py::array_t<float, py::array::forcecast | py::array::c_style> w3,
                    py::array_t<float, py::array::forcecast | py::array::c_style> ws) {
    check(hh_three_body_load(tb_, w3.data(), w3.size(), ws.data(), ws.size()));
  }
  void set_three_body_epsilon(float e) { check(hh_three_body_set_epsilon(tb_, e)); }
  void set_precision(const std::string &p) {
    int v = p == "fp32" ? HH_PREC_FP32 : (p == "fp16" ? HH_PREC_FP16 : -1);
    if (v < 0) throw std::invalid_argument("precision must be 'fp32' or 'fp16'");
    check(hh_three_body_set_precision(tb_, v));
  }
  py::array_t<float> velocity(py::array_t<float, py::array::forcecast> coord,
                              py::array_t<float, py::array::forcecast> force) {
    auto x = rows3(coord, "coord");
    auto f = rows3(force, "force");
    const py::ssize_t n = coord.shape(0);
    if (force.shape(0) != n) throw std::invalid_argument("coord and force sizes differ");
    py::array_t<float> out({n, (py::ssize_t)3});
    check(hh_three_body_velocity(tb_, x.data(), f.data(), (int64_t)n, out.mutable_data()));
    return out;
  }
  py::dict last_profile() {
    int64_t e[2];
    double ms[2];
    check(hh_three_body_last_profile(tb_, e, ms));
    py::dict d;
    d["three_body_edges"] = e[0];
    d["self_edges"] = e[1];
    d["three_body_ms"] = ms[0];
    d["self_ms"] = ms[1];
    return d;
  }

 private:
  hh_three_body *tb_ = nullptr;
};

// ---------------------------------------------------------------------------
// Time integrators (include/TimeIntegrator.hpp).  Positions are float
// triples updated with float arithmetic in the reference's order.  The
// reference never sets angular velocities (TimeIntegrator.hpp:137 is
// commented out), so orientations do not evolve: the Euler angles passed to
// the velocity callback stay those given to initialize() (the reference's
// zero-rate quaternion update yields NaN there, Quaternion.hpp:62-70).
// No VTK writer: without set_output_func no output is written.
// ---------------------------------------------------------------------------
struct V3 {
  float v[3] = {0, 0, 0};
};

class Integrator {
 public:
  virtual ~Integrator() {
    if (stepper_) hh_stepper_destroy(stepper_);
  }
  void set_time_step(float dt) { dt_ = dt; }
  void set_final_time(float t) { tf_ = t; }
  virtual void set_num_rigid_body(size_t n) {
    n_ = n;
    pos_.assign(n, V3());
    pos0_.assign(n, V3());
    vel_.assign(n, V3());
    off_.assign(n, V3());
    ang_.assign(3 * n, 0.0f);
  }
  void set_velocity_func(std::function<py::array_t<float>(float, py::array_t<float>)> f) { vf_ = f; }
  void set_output_func(std::function<void(py::array_t<float>, py::array_t<float>)> f) {
    of_ = f;
    custom_out_ = true;
  }
  void set_x_lim(py::list l) { set_lim(0, l); }
  void set_y_lim(py::list l) { set_lim(1, l); }
  void set_z_lim(py::list l) { set_lim(2, l); }
  void set_output_step(int s) { out_step_ = s; }
  // device-resident runs: stop after this many accepted steps (0: run to the
  // final time; hh_stepper_run's max_steps)
  void set_max_steps(int64_t s) {
    if (s < 0) throw std::runtime_error("max_steps must be >= 0");
    max_steps_ = s;
  }
  void initialize(py::array_t<float, py::array::forcecast> init) {
    if (init.ndim() != 2) throw std::runtime_error("Number of dimensions of input positions must be two");
    if ((size_t)init.shape(0) != n_) throw std::runtime_error("Inconsistent number of rigid bodies");
    auto d = init.unchecked<2>();
    for (size_t i = 0; i < n_; i++) {
      for (int j = 0; j < 3; j++) pos0_[i].v[j] = pos_[i].v[j] = d(i, j);
      for (int j = 0; j < 3; j++) ang_[3 * i + j] = init.shape(1) >= 6 ? d(i, 3 + j) : 0.0f;
    }
  }
  int func_count() const { return count_; }

  // Device-resident velocity: every stage evaluates `forcing` on the device,
  // updates the model's coordinates and applies M(x) on the device; no
  // Python callback and no host copy of the particle arrays per stage.
  void set_device_velocity(Model &model, Forcing &forcing, ThreeBody *three_body) {
    if ((size_t)model.size() != n_ && n_ != 0)
      throw std::runtime_error("model and integrator sizes differ");
    dm_ = &model;
    df_ = &forcing;
    dt3_ = three_body;
    if (stepper_) {
      hh_stepper_destroy(stepper_);
      stepper_ = nullptr;
    }
  }
  void set_graphs(bool flag) { graphs_ = flag; }
  // kept tree rebuilt every `every` accepted steps (0: never)
  void set_tree_rebuild(int64_t every) {
    if (every < 0) throw std::runtime_error("every must be >= 0");
    rebuild_every_ = every;
  }
  py::array_t<float> positions() const {
    py::array_t<float> p({(py::ssize_t)n_, (py::ssize_t)3});
    auto a = p.mutable_unchecked<2>();
    for (size_t i = 0; i < n_; i++)
      for (int j = 0; j < 3; j++) a(i, j) = pos_[i].v[j];
    return p;
  }
  py::dict device_stats() const {
    py::dict d;
    d["velocity_evaluations"] = stats_[0];
    d["accepted"] = stats_[1];
    d["rejected"] = stats_[2];
    d["whole_stage_graphs"] = stats_[3];
    return d;
  }

 protected:
  void set_lim(int d, py::list l) {
    lim_[0][d] = py::cast<float>(l[0]);
    lim_[1][d] = py::cast<float>(l[1]);
    periodic_ = true;
  }
  // TimeIntegrator::VelocityUpdate (TimeIntegrator.hpp:92-141)
  void velocity(float t, std::vector<V3> &out) {
    py::array_t<float> in({(py::ssize_t)n_, (py::ssize_t)6});
    auto a = in.mutable_unchecked<2>();
    for (size_t i = 0; i < n_; i++) {
      for (int j = 0; j < 3; j++) {
        if (periodic_) {
          const float x = pos_[i].v[j];
          off_[i].v[j] = x > lim_[1][j] ? lim_[0][j] - lim_[1][j]
                                        : (x < lim_[0][j] ? lim_[1][j] - lim_[0][j] : 0.0f);
          a(i, j) = x + off_[i].v[j];
        } else {
          a(i, j) = pos_[i].v[j];
        }
        a(i, 3 + j) = ang_[3 * i + j];
      }
    }
    py::array_t<float, py::array::forcecast> r = vf_(t, in);
    if (r.ndim() != 2 || (size_t)r.shape(0) != n_ || r.shape(1) < 3)
      throw std::runtime_error("velocity callback must return an (N,3) array");
    auto rv = r.unchecked<2>();
    for (size_t i = 0; i < n_; i++)
      for (int j = 0; j < 3; j++) out[i].v[j] = rv(i, j);
  }
  void output() {
    if (!custom_out_) return;
    py::array_t<float> p({(py::ssize_t)n_, (py::ssize_t)3}), v({(py::ssize_t)n_, (py::ssize_t)3});
    auto pa = p.mutable_unchecked<2>();
    auto va = v.mutable_unchecked<2>();
    for (size_t i = 0; i < n_; i++)
      for (int j = 0; j < 3; j++) pa(i, j) = pos_[i].v[j], va(i, j) = vel_[i].v[j];
    of_(p, v);
  }
  void wrap_offsets() {
    for (size_t i = 0; i < n_; i++)
      for (int j = 0; j < 3; j++) {
        const float x = pos_[i].v[j];
        off_[i].v[j] = x > lim_[1][j] ? lim_[0][j] - lim_[1][j]
                                      : (x < lim_[0][j] ? lim_[1][j] - lim_[0][j] : 0.0f);
      }
  }

  size_t n_ = 0;
  double dt_ = 1.0, tf_ = 10.0;
  int out_step_ = 1, count_ = 0;
  bool periodic_ = false, custom_out_ = false;
  int64_t max_steps_ = 0;
  float lim_[2][3] = {{-1, -1, -1}, {1, 1, 1}};
  std::vector<V3> pos_, pos0_, vel_, off_;
  std::vector<float> ang_;
  std::function<py::array_t<float>(float, py::array_t<float>)> vf_;
  std::function<void(py::array_t<float>, py::array_t<float>)> of_;

  // ---- device-resident mode (set_device_velocity) ----
  Model *dm_ = nullptr;
  Forcing *df_ = nullptr;
  ThreeBody *dt3_ = nullptr;
  hh_stepper *stepper_ = nullptr;
  bool graphs_ = true;
  int64_t rebuild_every_ = 0;
  int64_t stats_[4] = {0, 0, 0, 0};

  bool device_mode() const { return dm_ != nullptr; }

  static void device_output(void *user, int64_t, double, const float *, const float *) {
    auto *self = static_cast<Integrator *>(user);
    py::array_t<float> p({(py::ssize_t)self->n_, (py::ssize_t)3}),
        v({(py::ssize_t)self->n_, (py::ssize_t)3});
    check(hh_stepper_state(self->stepper_, p.mutable_data(), v.mutable_data(), nullptr, nullptr));
    self->of_(p, v);
  }

  // Runs the integration on the device from the current positions (a run
  // starts at t = 0, as the host loops do).
  void run_device(int method, double threshold) {
    if ((size_t)dm_->size() != n_) throw std::runtime_error("model and integrator sizes differ");
    if (!stepper_) check(hh_stepper_create(dm_->handle(), df_->handle(), method, &stepper_));
    check(hh_stepper_set_three_body(stepper_, dt3_ ? dt3_->handle() : nullptr));
    check(hh_stepper_set(stepper_, dt_, tf_, threshold));
    if (periodic_)
      for (int d = 0; d < 3; d++) check(hh_stepper_set_lim(stepper_, d, lim_[0][d], lim_[1][d]));
    check(hh_stepper_set_graphs(stepper_, graphs_));
    check(hh_stepper_set_tree_rebuild(stepper_, rebuild_every_));
    std::vector<float> x(3 * n_);
    for (size_t i = 0; i < n_; i++)
      for (int j = 0; j < 3; j++) x[3 * i + j] = pos_[i].v[j];
    check(hh_stepper_initialize(stepper_, x.data(), (int64_t)n_));
    check(hh_stepper_set_output(stepper_, out_step_, custom_out_ ? &device_output : nullptr, this));
    check(hh_stepper_run(stepper_, max_steps_));
    std::vector<float> v(3 * n_);
    check(hh_stepper_state(stepper_, x.data(), v.data(), nullptr, stats_));
    for (size_t i = 0; i < n_; i++)
      for (int j = 0; j < 3; j++) pos_[i].v[j] = x[3 * i + j], vel_[i].v[j] = v[3 * i + j];
    count_ = (int)stats_[1];
  }
};

// ExplicitEuler::Run (TimeIntegrator.hpp:431-488).
class Euler : public Integrator {
 public:
  void run() {
    if (device_mode()) return run_device(HH_STEP_EULER, 0.0);
    count_ = 0;
    double t = 0.0;
    const double h0 = dt_;
    while (t < tf_ - 1e-5) {
      const double dt = std::min(h0, tf_ - t);
      velocity((float)(t + dt), vel_);
      const float fdt = (float)dt;
      for (size_t i = 0; i < n_; i++)
        for (int j = 0; j < 3; j++) pos_[i].v[j] = pos_[i].v[j] + fdt * vel_[i].v[j];
      if (periodic_) wrap_offsets();
      t += dt;
      if (count_ % out_step_ == 0) output();
      if (periodic_)
        for (size_t i = 0; i < n_; i++)
          for (int j = 0; j < 3; j++) pos_[i].v[j] = pos_[i].v[j] + off_[i].v[j];
      count_++;
    }
  }
};

// ExplicitRk4::Run (TimeIntegrator.hpp:602-860): Dormand-Prince 5(4) with
// FSAL.  K1 starts at zero (the first call's velocity is not stored in K1),
// a rejected step shrinks dt by max(0.8 (err/thr)^-0.2, 0.1) (>= 1e-6), and
// the step size is reset to the nominal one at the start of every step.
class Rk4 : public Integrator {
 public:
  void set_num_rigid_body(size_t n) override {
    Integrator::set_num_rigid_body(n);
    for (auto &k : K_) k.assign(n, V3());
  }
  void set_threshold(double t) { thr_ = t; }
  void run() {
    const float a21 = 1.0f / 5.0f;
    const float a31 = 3.0f / 40.0f, a32 = 9.0f / 40.0f;
    const float a41 = 44.0f / 45.0f, a42 = -56.0f / 15.0f, a43 = 32.0f / 9.0f;
    const float a51 = 19372.0f / 6561.0f, a52 = -25360.0f / 2187.0f, a53 = 64448.0f / 6561.0f,
                a54 = -212.0f / 729.0f;
    const float a61 = 9017.0f / 3168.0f, a62 = -355.0f / 33.0f, a63 = 46732.0f / 5247.0f,
                a64 = 49.0f / 176.0f, a65 = -5103.0f / 18656.0f;
    const float b1 = 35.0f / 384.0f, b3 = 500.0f / 1113.0f, b4 = 125.0f / 192.0f,
                b5 = -2187.0f / 6784.0f, b6 = 11.0f / 84.0f;
    const float dc1 = b1 - 5179.0f / 57600.0f, dc3 = b3 - 7571.0f / 16695.0f,
                dc4 = b4 - 393.0f / 640.0f, dc5 = b5 - (-92097.0f / 339200.0f),
                dc6 = b6 - 187.0f / 2100.0f, dc7 = -1.0f / 40.0f;
    const float c[7] = {0.0f, 1.0f / 5.0f, 3.0f / 10.0f, 4.0f / 5.0f, 8.0f / 9.0f, 1.0f, 1.0f};
    if (device_mode()) return run_device(HH_STEP_DOPRI, thr_);
    const float A[6][5] = {{a21, 0, 0, 0, 0},       {a31, a32, 0, 0, 0},
                           {a41, a42, a43, 0, 0},   {a51, a52, a53, a54, 0},
                           {a61, a62, a63, a64, a65}, {b1, 0, b3, b4, b5}};
    count_ = 0;
    double t = 0.0;
    const double h0 = dt_;
    velocity((float)t, vel_);
    while (t < tf_ - 1e-5) {
      double dt = std::min(h0, tf_ - t);
      for (size_t i = 0; i < n_; i++)
        for (int j = 0; j < 3; j++)
          pos0_[i].v[j] = periodic_ ? pos_[i].v[j] + off_[i].v[j] : pos_[i].v[j];
      double err = 100.0;
      while (err > thr_) {
        for (int s = 1; s < 7; s++) {
          // stage positions: x0 + sum_k K_k (a_sk dt), left to right
          for (size_t i = 0; i < n_; i++)
            for (int j = 0; j < 3; j++) {
              float x = pos0_[i].v[j];
              for (int k = 0; k < s && k < 5; k++) {
                if (s == 6 && k == 1) continue;  // b2 = 0
                x = x + (float)(A[s - 1][k] * dt) * K_[k][i].v[j];
              }
              if (s == 6) x = x + (float)(b6 * dt) * K_[5][i].v[j];
              pos_[i].v[j] = x;
            }
          velocity((float)(t + c[s] * dt), vel_);
          K_[s] = vel_;
        }
        err = 0.0;
        const float fdt = (float)dt;
        for (size_t i = 0; i < n_; i++) {
          float e2 = 0.0f;
          for (int j = 0; j < 3; j++) {
            float v = K_[0][i].v[j] * dc1 + K_[2][i].v[j] * dc3;
            v = v + K_[3][i].v[j] * dc4;
            v = v + K_[4][i].v[j] * dc5;
            v = v + K_[5][i].v[j] * dc6;
            v = v + K_[6][i].v[j] * dc7;
            v = fdt * v;
            e2 += v * v;
          }
          err += e2;
        }
        err = std::sqrt(err);
        if (err > thr_) {
          dt = dt * std::max(0.8 * std::pow(err / thr_, -0.2), 0.1);
          dt = std::max(dt, 1e-6);
        }
      }
      t += dt;
      if (count_ % out_step_ == 0) output();
      K_[0] = K_[6];
      count_++;
    }
  }

 private:
  double thr_ = 1e-3;
  std::vector<V3> K_[7];
};

}  // namespace

PYBIND11_MODULE(hignn, m) {
  m.doc() = "H-HIGNN mobility matvec on MI355X (gfx950)";
  m.def("Init", []() { check(hh_init()); }, "Select the HIP device");
  m.def("Finalize", []() { check(hh_finalize()); }, "Synchronise the device");

  py::class_<Model>(m, "HignnModel")
      .def(py::init<py::array_t<float, py::array::forcecast>, int>())
      .def("load_two_body_model", &Model::load_two_body_model)
      .def("set_tree_reuse", &Model::set_tree_reuse)
      .def("set_tree_build", &Model::set_tree_build)
      .def("load_two_body_weights", &Model::load_two_body_weights)
      .def("load_three_body_model", &Model::load_three_body_model)
      .def("set_epsilon", &Model::set_epsilon)
      .def("set_eta", &Model::set_eta)
      .def("set_max_iter", &Model::set_max_iter)
      .def("set_mat_pool_size_factor", &Model::set_mat_pool_size_factor)
      .def("set_max_far_dot_work_node_size", &Model::set_max_far_dot_work_node_size)
      .def("set_max_relative_coord", &Model::set_max_relative_coord)
      .def("set_post_check_flag", &Model::set_post_check_flag)
      .def("set_use_symmetry_flag", &Model::set_use_symmetry_flag)
      .def("set_max_far_field_distance", &Model::set_max_far_field_distance)
      .def("set_precision", &Model::set_precision)
      .def("update_coord", &Model::update_coord)
      .def("dot", &Model::dot)
      .def("dense_dot", &Model::dense_dot)
      .def("dot_device", &Model::dot_device, py::arg("u_ptr"), py::arg("f_ptr"),
           py::arg("stream") = 0)
      .def("tune_far_field", &Model::tune_far_field, py::arg("f_ptr"), py::arg("stream") = 0)
      .def("tune_near", &Model::tune_near, py::arg("f_ptr"), py::arg("stream") = 0)
      .def("set_near_shape", &Model::set_near_shape)
      .def("set_near_workgroups", &Model::set_near_workgroups)
      .def("tree_export", &Model::tree_export)
      .def("far_ranks", &Model::far_ranks)
      .def("check_fp64", &Model::check_fp64, py::arg("f"), py::arg("targets"),
           py::arg("convention") = 0)
      .def("last_timings", &Model::last_timings)
      .def("last_profile", &Model::last_profile);

  py::class_<Euler>(m, "ExplicitEuler")
      .def(py::init())
      .def("set_time_step", &Euler::set_time_step)
      .def("set_final_time", &Euler::set_final_time)
      .def("set_num_rigid_body", &Euler::set_num_rigid_body)
      .def("set_velocity_func", &Euler::set_velocity_func)
      .def("set_output_func", &Euler::set_output_func)
      .def("set_x_lim", &Euler::set_x_lim)
      .def("set_y_lim", &Euler::set_y_lim)
      .def("set_z_lim", &Euler::set_z_lim)
      .def("set_output_step", &Euler::set_output_step)
      .def("set_max_steps", &Euler::set_max_steps)
      .def("initialize", &Euler::initialize)
      .def("run", &Euler::run)
      .def("func_count", &Euler::func_count)
      .def("set_device_velocity", &Euler::set_device_velocity, py::arg("model"),
           py::arg("forcing"), py::arg("three_body") = nullptr, py::keep_alive<1, 2>(),
           py::keep_alive<1, 3>(), py::keep_alive<1, 4>())
      .def("set_graphs", &Euler::set_graphs)
      .def("set_tree_rebuild", &Euler::set_tree_rebuild, py::arg("every"))
      .def("positions", &Euler::positions)
      .def("device_stats", &Euler::device_stats);

  py::class_<Rk4>(m, "ExplicitRk4")
      .def(py::init())
      .def("set_time_step", &Rk4::set_time_step)
      .def("set_threshold", &Rk4::set_threshold)
      .def("set_final_time", &Rk4::set_final_time)
      .def("set_num_rigid_body", &Rk4::set_num_rigid_body)
      .def("set_velocity_func", &Rk4::set_velocity_func)
      .def("set_output_func", &Rk4::set_output_func)
      .def("set_x_lim", &Rk4::set_x_lim)
      .def("set_y_lim", &Rk4::set_y_lim)
      .def("set_z_lim", &Rk4::set_z_lim)
      .def("set_output_step", &Rk4::set_output_step)
      .def("set_max_steps", &Rk4::set_max_steps)
      .def("initialize", &Rk4::initialize)
      .def("run", &Rk4::run)
      .def("func_count", &Rk4::func_count)
      .def("set_device_velocity", &Rk4::set_device_velocity, py::arg("model"),
           py::arg("forcing"), py::arg("three_body") = nullptr, py::keep_alive<1, 2>(),
           py::keep_alive<1, 3>(), py::keep_alive<1, 4>())
      .def("set_graphs", &Rk4::set_graphs)
      .def("set_tree_rebuild", &Rk4::set_tree_rebuild, py::arg("every"))
      .def("positions", &Rk4::positions)
      .def("device_stats", &Rk4::device_stats);

  py::class_<Forcing>(m, "DeviceForcing")
      .def(py::init())
      .def("set_body_force", &Forcing::set_body_force)
      .def("set_filaments", &Forcing::set_filaments, py::arg("n_filament"), py::arg("n_chain"),
           py::arg("rest_length"), py::arg("k_t"), py::arg("k_b"))
      .def("set_bonds", &Forcing::set_bonds, py::arg("n_filament"), py::arg("n_chain"),
           py::arg("D") = 1.0, py::arg("alpha") = 1.0, py::arg("r0") = 2.2)
      .def("set_two_body_epsilon", &Forcing::set_two_body_epsilon)
      .def("set_periodic", &Forcing::set_periodic, py::arg("is_periodic") = false)
      .def("set_domain", &Forcing::set_domain)
      .def("eval", &Forcing::eval);

  py::class_<Neighbors>(m, "NeighborLists")
      .def(py::init())
      .def("set_periodic", &Neighbors::set_periodic, py::arg("is_periodic") = false)
      .def("set_domain", &Neighbors::set_domain)
      .def("update_coord", &Neighbors::update_coord)
      .def("set_two_body_epsilon", &Neighbors::set_two_body_epsilon)
      .def("set_three_body_epsilon", &Neighbors::set_three_body_epsilon)
      .def("get_two_body_edge_info", &Neighbors::two_body)
      .def("get_two_body_edge_info_by_index", &Neighbors::two_body_by_index)
      .def("get_three_body_edge_info", &Neighbors::three_body)
      .def("get_three_body_edge_info_by_index", &Neighbors::three_body_by_index)
      .def("get_three_body_edge_self_info", &Neighbors::three_body_self)
      .def("get_three_body_edge_self_info_by_index", &Neighbors::three_body_self_by_index)
      .def("build_two_body_info", &Neighbors::build_two_body_info)
      .def("build_three_body_info", &Neighbors::build_three_body_info);

  py::class_<ThreeBody>(m, "ThreeBodyModel")
      .def(py::init())
      .def("load_model", &ThreeBody::load_model)
      .def("load_weights", &ThreeBody::load_weights)
      .def("set_three_body_epsilon", &ThreeBody::set_three_body_epsilon)
      .def("set_precision", &ThreeBody::set_precision)
      .def("velocity", &ThreeBody::velocity)
      .def("last_profile", &ThreeBody::last_profile);

  py::class_<PotentialForceCls, Forcing>(m, "PotentialForce")
      .def(py::init())
      .def("get_potential_force", &Forcing::eval);
}
