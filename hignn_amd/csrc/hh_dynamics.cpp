// Device forcing programs and device-resident integrators (C ABI of
// include/hh_dynamics.h).
//
// An integrator stage runs, on one stream: the stage combination of the
// positions and the forcing (captured once into a HIP graph per stage and
// replayed), then update_coord and u = M(x) f of the model.  The host only
// sees the adaptive step's error norm (one double per attempt).  The float
// operations and their order follow the host integrators of the Python
// module, which follow include/TimeIntegrator.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/hh_dynamics.h"
#include "../../include/hh_three_body.h"
#include "hh_cells.h"
#include "hh_forcing.h"
#include "hh_host.h"

using namespace hh;

struct hh_forcing {
  float b[3] = {0, 0, 0};
  int64_t n_fil = 0, n_chain = 0;
  float rest = 0, k_t = 0;
  double k_b = 0;
  int64_t bond_fil = 0, bond_chain = 0;
  double bond_D = 0, bond_alpha = 0, bond_r0 = 0;
  float eps = 0;
  bool periodic = false, has_domain = false;
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  int version = 0;  // bumped by every setter (integrators re-capture their graphs)
  CellBufs cells;
  DBuf xh, fh;

  // Reserves the buffers for n particles and fills the launch descriptor.
  ForcingLaunch prepare(const float *x, float *f, int64_t n) {
    if (eps > 0 && periodic && !has_domain)
      throw HHError(HH_ERR_STATE, "periodic Morse force without a domain (set_domain)");
    if (n_fil > 0 && n_fil * n_chain > n)
      throw HHError(HH_ERR_ARG, "filaments cover more particles than given");
    if (bond_fil > 0 && bond_fil * bond_chain > n)
      throw HHError(HH_ERR_ARG, "bond chains cover more particles than given");
    ForcingLaunch L{};
    L.x = x;
    L.f = f;
    L.n = n;
    for (int j = 0; j < 3; j++) L.b[j] = b[j], L.lo[j] = lo[j], L.hi[j] = hi[j];
    L.n_fil = n_fil;
    L.n_chain = n_chain;
    L.rest = rest;
    L.k_t = k_t;
    L.k_b = k_b;
    L.bond_fil = bond_fil;
    L.bond_chain = bond_chain;
    L.bond_D = bond_D;
    L.bond_alpha = bond_alpha;
    L.bond_r0 = bond_r0;
    L.morse = eps > 0 ? 1 : 0;
    L.periodic = periodic ? 1 : 0;
    L.eps = eps;
    L.margin = eps;
    if (!L.morse || n <= 0) return L;
    const std::array<void *, 11> before = cells.pointers();
    cells.setup(L, n);
    if (cells.pointers() != before) version++;  // moved buffers invalidate captured graphs
    return L;
  }

  ~hh_forcing() {
    cells.release();
    xh.release();
    fh.release();
  }
};

struct hh_stepper {
  hh_model *m = nullptr;
  hh_forcing *fc = nullptr;
  hh_three_body *tb = nullptr;  // optional three-body / self corrections
  int method = HH_STEP_EULER;
  int64_t n = 0;
  double dt = 1.0, tf = 10.0, thr = 1e-3;
  bool periodic = false, use_graphs = true, initialized = false, started = false;
  Lim lim{{-1, -1, -1}, {1, 1, 1}};
  hipStream_t st = nullptr;
  DBuf X, X0, XIN, OFF, F, V, K[7], coef, part, err;
  PBuf hcoef, herr;
  // graphs: [0] Euler input + forcing, [1..6] DOPRI stages + forcing
  hipGraphExec_t g[7] = {};
  int g_version = -1;
  // whole-stage graphs (stage, forcing, coordinate gather and matvec) while
  // the model's matvec is capturable; valid for one model signature
  hipGraphExec_t fg[7] = {};
  int fg_version = -1;
  uint64_t fg_sig = 0;
  int64_t full_launches = 0;
  double t = 0;
  int64_t count = 0, evals = 0, rejected = 0;
  int out_step = 1;
  hh_output_fn out_fn = nullptr;
  void *out_user = nullptr;

  ~hh_stepper() {
    drop_graphs();
    for (DBuf *d : {&X, &X0, &XIN, &OFF, &F, &V, &coef, &part, &err}) d->release();
    for (DBuf &k : K) k.release();
    hcoef.release();
    herr.release();
    if (st) (void)hipStreamDestroy(st);
  }

  void drop_graphs() {
    for (auto &e : g)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
    g_version = -1;
    drop_full();
  }

  void drop_full() {
    for (auto &e : fg)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
    fg_version = -1;
  }

  // Stage descriptor: row 0 = Euler input (x0 = X), rows 1..6 = DOPRI stages.
  StageLaunch stage(int row) {
    StageLaunch S{};
    S.n = n;
    S.periodic = periodic ? 1 : 0;
    for (int j = 0; j < 3; j++) S.lo[j] = lim.lo[j], S.hi[j] = lim.hi[j];
    S.coef = coef.as<float>();
    S.row = row;
    S.xin = XIN.as<float>();
    if (row == 0) {
      S.x0 = X.as<float>();
      S.nterms = 0;
      S.xs = nullptr;
      return S;
    }
    S.x0 = X0.as<float>();
    S.xs = X.as<float>();
    int q = 0;
    if (row < 6) {
      for (int k = 0; k < row; k++) S.k[q++] = K[k].as<float>();
    } else {
      for (int k : {0, 2, 3, 4, 5}) S.k[q++] = K[k].as<float>();
    }
    S.nterms = q;
    return S;
  }

  void enqueue_prep(int row) {
    HIPCHK(launch_stage(stage(row), st));
    const ForcingLaunch L = fc->prepare(XIN.as<float>(), F.as<float>(), n);
    HIPCHK(launch_forcing(L, st));
  }

  template <class Enqueue>
  void capture(hipGraphExec_t &exec, Enqueue &&enqueue) {
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      enqueue();
    } catch (...) {
      (void)hipStreamEndCapture(st, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    HIPCHK(hipStreamEndCapture(st, &graph));
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
  }

  void run_prep(int row) {
    if (!use_graphs) {
      enqueue_prep(row);
      return;
    }
    // reserve every forcing buffer before capture (no allocation inside)
    (void)fc->prepare(XIN.as<float>(), F.as<float>(), n);
    if (g_version != fc->version) {
      drop_graphs();
      g_version = fc->version;
    }
    if (!g[row]) capture(g[row], [&]() { enqueue_prep(row); });
    HIPCHK(hipGraphLaunch(g[row], st));
  }

  void model_call(int rc) {
    if (rc != HH_OK) throw HHError(HH_ERR_STATE, std::string("model: ") + hh_last_error());
  }

  // The whole stage as one graph when the matvec allows it; false otherwise.
  bool run_full(int row, DBuf &out) {
    if (!use_graphs || tb) return false;  // the three-body term waits on the host
    int cap = 0;
    uint64_t sig = 0;
    model_call(hh_model_dot_capturable(m, &cap, &sig));
    if (!cap) return false;
    (void)fc->prepare(XIN.as<float>(), F.as<float>(), n);
    if (fg_version != fc->version || fg_sig != sig) {
      drop_full();
      fg_version = fc->version;
      fg_sig = sig;
    }
    if (!fg[row])
      capture(fg[row], [&]() {
        enqueue_prep(row);
        model_call(hh_model_update_coord_device(m, XIN.as<float>(), n, st));
        model_call(hh_model_dot_enqueue(m, out.as<float>(), F.as<float>(), n, st));
      });
    HIPCHK(hipGraphLaunch(fg[row], st));
    full_launches++;
    return true;
  }

  // Velocity of a stage: forcing at the stage input, update_coord, u = M f.
  void velocity(int row, DBuf &out) {
    const bool rebuild = pending_rebuild;
    pending_rebuild = false;
    if (!rebuild && run_full(row, out)) {
      evals++;
      return;
    }
    run_prep(row);
    if (rebuild) {
      model_call(hh_model_set_tree_reuse(m, 0));
      const int rc = hh_model_update_coord_device(m, XIN.as<float>(), n, st);
      model_call(hh_model_set_tree_reuse(m, 1));
      model_call(rc);
      rebuilds++;
    } else {
      model_call(hh_model_update_coord_device(m, XIN.as<float>(), n, st));
    }
    model_call(hh_model_dot_device(m, out.as<float>(), F.as<float>(), n, st));
    if (tb && hh_three_body_velocity_device(tb, XIN.as<float>(), F.as<float>(), n, out.as<float>(), 1,
                                            st) != HH_OK)
      throw HHError(HH_ERR_STATE, std::string("three-body: ") + hh_last_error());
    evals++;
  }

  void output(const DBuf &v) {
    if (!out_fn || count % out_step != 0) return;
    HIPCHK(hipStreamSynchronize(st));
    out_fn(out_user, count, t, X.as<float>(), v.as<float>());
  }

  // With the model's tree kept across stages (hh_model_set_tree_reuse), a
  // full rebuild at the start of every rebuild_every-th accepted step
  // (0: never): the positions have moved, and admissibility is re-checked.
  // The rebuild uses the first stage's input positions, as the first build
  // of the run does.
  int64_t rebuild_every = 0, rebuilds = 0;
  // HH_STEPPER_LOG=1: one line on stderr per adaptive attempt (the reference
  // prints each step's time and dt, include/TimeIntegrator.hpp:612-615)
  const bool log_steps = std::getenv("HH_STEPPER_LOG") && std::atoi(std::getenv("HH_STEPPER_LOG")) != 0;
  bool pending_rebuild = false;
  void maybe_rebuild() {
    if (rebuild_every > 0 && count > 0 && count % rebuild_every == 0) pending_rebuild = true;
  }

  // ExplicitEuler::Run (TimeIntegrator.hpp:431-488)
  void run_euler(int64_t max_steps) {
    const double h0 = dt;
    int64_t taken = 0;
    while (t < tf - 1e-5 && (max_steps <= 0 || taken < max_steps)) {
      const double h = std::min(h0, tf - t);
      maybe_rebuild();
      velocity(0, V);
      HIPCHK(launch_set_f32(coef.as<float>() + 7 * 8, (float)h, st));
      HIPCHK(launch_euler_update(X.as<float>(), V.as<float>(), OFF.as<float>(), n,
                                 coef.as<float>() + 7 * 8, periodic ? 1 : 0, lim, st));
      t += h;
      output(V);
      if (periodic) HIPCHK(launch_add(X.as<float>(), OFF.as<float>(), 3 * n, st));
      count++;
      taken++;
    }
    HIPCHK(hipStreamSynchronize(st));
  }

  // ExplicitRk4::Run (TimeIntegrator.hpp:602-860): Dormand-Prince 5(4), FSAL.
  void run_dopri(int64_t max_steps) {
    // the tableau in float, as the host integrator defines it
    const float b1 = 35.0f / 384.0f, b3 = 500.0f / 1113.0f, b4 = 125.0f / 192.0f,
                b5 = -2187.0f / 6784.0f, b6 = 11.0f / 84.0f;
    const float Af[6][5] = {
        {1.0f / 5.0f, 0, 0, 0, 0},
        {3.0f / 40.0f, 9.0f / 40.0f, 0, 0, 0},
        {44.0f / 45.0f, -56.0f / 15.0f, 32.0f / 9.0f, 0, 0},
        {19372.0f / 6561.0f, -25360.0f / 2187.0f, 64448.0f / 6561.0f, -212.0f / 729.0f, 0},
        {9017.0f / 3168.0f, -355.0f / 33.0f, 46732.0f / 5247.0f, 49.0f / 176.0f,
         -5103.0f / 18656.0f},
        {b1, 0, b3, b4, b5}};
    ErrLaunch E{};
    E.dc[0] = b1 - 5179.0f / 57600.0f;
    E.dc[2] = b3 - 7571.0f / 16695.0f;
    E.dc[3] = b4 - 393.0f / 640.0f;
    E.dc[4] = b5 - (-92097.0f / 339200.0f);
    E.dc[5] = b6 - 187.0f / 2100.0f;
    E.dc[6] = -1.0f / 40.0f;
    for (int k = 0; k < 7; k++) E.k[k] = K[k].as<float>();
    E.fdt = coef.as<float>() + 7 * 8;
    E.part = part.as<double>();
    E.n = n;

    if (!started) {
      velocity(0, V);  // the first call's velocity is not kept in K1
      started = true;
    }
    const double h0 = dt;
    int64_t taken = 0;
    while (t < tf - 1e-5 && (max_steps <= 0 || taken < max_steps)) {
      double h = std::min(h0, tf - t);
      maybe_rebuild();
      // x0 = x (+ periodic offset of x)
      StageLaunch S0 = stage(0);
      S0.xin = X0.as<float>();
      HIPCHK(launch_stage(S0, st));
      double e = 100.0;
      while (e > thr) {
        float *c = hcoef.as<float>();
        for (int s = 1; s <= 6; s++) {
          int q = 0;
          if (s < 6) {
            for (int k = 0; k < s; k++) c[s * 8 + q++] = (float)((double)Af[s - 1][k] * h);
          } else {
            for (int k : {0, 2, 3, 4}) c[s * 8 + q++] = (float)((double)Af[5][k] * h);
            c[s * 8 + q++] = (float)((double)b6 * h);
          }
        }
        c[7 * 8] = (float)h;
        HIPCHK(hipMemcpyAsync(coef.as<float>(), c, 8 * 8 * 4, hipMemcpyHostToDevice, st));
        for (int s = 1; s <= 6; s++) velocity(s, K[s]);
        HIPCHK(launch_dopri_err(E, err.as<double>(), st));
        HIPCHK(hipMemcpyAsync(herr.p, err.p, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        e = std::sqrt(*herr.as<double>());
        if (log_steps) {
          std::fprintf(stderr, "[stepper] t %.6f h %.4e err %.4e %s (%lld evaluations)\n", t, h, e,
                       e > thr ? "rejected" : "accepted", (long long)evals);
          std::fflush(stderr);
        }
        if (e > thr) {
          h = h * std::max(0.8 * std::pow(e / thr, -0.2), 0.1);
          h = std::max(h, 1e-6);
          rejected++;
        }
      }
      t += h;
      HIPCHK(hipMemcpyAsync(V.p, K[6].p, 12 * (size_t)n, hipMemcpyDeviceToDevice, st));
      output(V);
      HIPCHK(hipMemcpyAsync(K[0].p, K[6].p, 12 * (size_t)n, hipMemcpyDeviceToDevice, st));
      count++;
      taken++;
    }
    HIPCHK(hipStreamSynchronize(st));
  }
};

namespace {
hh_forcing *need(hh_forcing *fc) {
  if (!fc) throw HHError(HH_ERR_ARG, "NULL forcing");
  return fc;
}
hh_stepper *need(hh_stepper *s) {
  if (!s) throw HHError(HH_ERR_ARG, "NULL stepper");
  return s;
}
}  // namespace

extern "C" {

int hh_forcing_create(hh_forcing **out) {
  return guarded([&]() {
    if (!out) throw HHError(HH_ERR_ARG, "NULL argument");
    ensure_device();
    *out = new hh_forcing();
  });
}

void hh_forcing_destroy(hh_forcing *fc) { delete fc; }

int hh_forcing_set_body(hh_forcing *fc, const float b[3]) {
  return guarded([&]() {
    if (!b) throw HHError(HH_ERR_ARG, "NULL argument");
    for (int j = 0; j < 3; j++) need(fc)->b[j] = b[j];
    fc->version++;
  });
}

int hh_forcing_set_filaments(hh_forcing *fc, int64_t n_filament, int64_t n_chain,
                             float rest_length, float k_t, double k_b) {
  return guarded([&]() {
    need(fc);
    if (n_filament < 0 || (n_filament > 0 && n_chain < 1))
      throw HHError(HH_ERR_ARG, "invalid filament layout");
    fc->n_fil = n_filament;
    fc->n_chain = n_filament > 0 ? n_chain : 0;
    fc->rest = rest_length;
    fc->k_t = k_t;
    fc->k_b = k_b;
    fc->version++;
  });
}

int hh_forcing_set_bonds(hh_forcing *fc, int64_t n_filament, int64_t n_chain, double D,
                         double alpha, double r0) {
  return guarded([&]() {
    need(fc);
    if (n_filament < 0 || (n_filament > 0 && n_chain < 1))
      throw HHError(HH_ERR_ARG, "invalid bond-chain layout");
    if (n_filament > 0 && !(alpha > 0 && r0 > 0)) throw HHError(HH_ERR_ARG, "need alpha, r0 > 0");
    fc->bond_fil = n_filament;
    fc->bond_chain = n_filament > 0 ? n_chain : 0;
    fc->bond_D = D;
    fc->bond_alpha = alpha;
    fc->bond_r0 = r0;
    fc->version++;
  });
}

int hh_forcing_set_morse(hh_forcing *fc, float epsilon) {
  return guarded([&]() {
    need(fc)->eps = epsilon > 0 ? epsilon : 0.0f;
    fc->version++;
  });
}

int hh_forcing_set_periodic(hh_forcing *fc, int flag) {
  return guarded([&]() {
    need(fc)->periodic = flag != 0;
    fc->version++;
  });
}

int hh_forcing_set_domain(hh_forcing *fc, const float lo[3], const float hi[3]) {
  return guarded([&]() {
    if (!lo || !hi) throw HHError(HH_ERR_ARG, "NULL argument");
    need(fc);
    for (int j = 0; j < 3; j++) fc->lo[j] = lo[j], fc->hi[j] = hi[j];
    fc->has_domain = true;
    fc->periodic = true;
    fc->version++;
  });
}

int hh_forcing_eval_device(hh_forcing *fc, const float *x_dev, float *f_dev, int64_t n,
                           void *stream) {
  return guarded([&]() {
    if (!x_dev || !f_dev || n < 0) throw HHError(HH_ERR_ARG, "invalid argument");
    const ForcingLaunch L = need(fc)->prepare(x_dev, f_dev, n);
    HIPCHK(launch_forcing(L, (hipStream_t)stream));
  });
}

int hh_forcing_eval(hh_forcing *fc, const float *x, float *f, int64_t n) {
  return guarded([&]() {
    if (!x || !f || n < 0) throw HHError(HH_ERR_ARG, "invalid argument");
    need(fc);
    if (n == 0) return;
    fc->xh.reserve(12 * (size_t)n);
    fc->fh.reserve(12 * (size_t)n);
    HIPCHK(hipMemcpy(fc->xh.p, x, 12 * (size_t)n, hipMemcpyHostToDevice));
    const ForcingLaunch L = fc->prepare(fc->xh.as<float>(), fc->fh.as<float>(), n);
    HIPCHK(launch_forcing(L, nullptr));
    HIPCHK(hipMemcpy(f, fc->fh.p, 12 * (size_t)n, hipMemcpyDeviceToHost));
  });
}

int hh_stepper_create(hh_model *m, hh_forcing *fc, int method, hh_stepper **out) {
  return guarded([&]() {
    if (!m || !fc || !out) throw HHError(HH_ERR_ARG, "NULL argument");
    if (method != HH_STEP_EULER && method != HH_STEP_DOPRI)
      throw HHError(HH_ERR_ARG, "unknown integration method");
    ensure_device();
    auto *s = new hh_stepper();
    s->m = m;
    s->fc = fc;
    s->method = method;
    hipError_t e = hipStreamCreate(&s->st);
    if (e != hipSuccess) {
      delete s;
      throw HHError(HH_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = s;
  });
}

void hh_stepper_destroy(hh_stepper *s) { delete s; }

int hh_stepper_set(hh_stepper *s, double time_step, double final_time, double threshold) {
  return guarded([&]() {
    if (!(time_step > 0)) throw HHError(HH_ERR_ARG, "time step must be positive");
    need(s)->dt = time_step;
    s->tf = final_time;
    s->thr = threshold;
  });
}

int hh_stepper_set_lim(hh_stepper *s, int axis, float lo, float hi) {
  return guarded([&]() {
    if (axis < 0 || axis > 2) throw HHError(HH_ERR_ARG, "axis must be 0, 1 or 2");
    need(s)->lim.lo[axis] = lo;
    s->lim.hi[axis] = hi;
    s->periodic = true;
    s->drop_graphs();
  });
}

int hh_stepper_set_three_body(hh_stepper *s, hh_three_body *tb) {
  return guarded([&]() { need(s)->tb = tb; });
}

int hh_stepper_set_tree_rebuild(hh_stepper *s, int64_t every) {
  return guarded([&]() {
    if (every < 0) throw HHError(HH_ERR_ARG, "every must be >= 0");
    need(s)->rebuild_every = every;
  });
}

int hh_stepper_set_graphs(hh_stepper *s, int flag) {
  return guarded([&]() {
    need(s)->use_graphs = flag != 0;
    s->drop_graphs();
  });
}

int hh_stepper_initialize(hh_stepper *s, const float *x, int64_t n) {
  return guarded([&]() {
    if (!x || n <= 0) throw HHError(HH_ERR_ARG, "invalid positions");
    need(s);
    if (s->n != n) {
      s->drop_graphs();
      s->n = n;
    }
    const size_t b = 12 * (size_t)n;
    for (DBuf *d : {&s->X, &s->X0, &s->XIN, &s->OFF, &s->F, &s->V}) d->reserve(b);
    for (DBuf &k : s->K) k.reserve(b);
    s->coef.reserve(8 * 8 * 4);
    s->part.reserve(8 * (size_t)err_partials_needed(n));
    s->err.reserve(8);
    s->hcoef.reserve(8 * 8 * 4);
    s->herr.reserve(8);
    HIPCHK(hipMemcpyAsync(s->X.p, x, b, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemsetAsync(s->OFF.p, 0, b, s->st));
    HIPCHK(hipMemsetAsync(s->V.p, 0, b, s->st));
    for (DBuf &k : s->K) HIPCHK(hipMemsetAsync(k.p, 0, b, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    s->t = 0;
    s->count = s->evals = s->rejected = s->full_launches = 0;
    s->started = false;
    s->initialized = true;
  });
}

int hh_stepper_set_output(hh_stepper *s, int output_step, hh_output_fn fn, void *user) {
  return guarded([&]() {
    if (output_step < 1) throw HHError(HH_ERR_ARG, "output step must be at least 1");
    need(s)->out_step = output_step;
    s->out_fn = fn;
    s->out_user = user;
  });
}

int hh_stepper_run(hh_stepper *s, int64_t max_steps) {
  return guarded([&]() {
    if (!need(s)->initialized) throw HHError(HH_ERR_STATE, "initialize has not run");
    if (s->method == HH_STEP_EULER)
      s->run_euler(max_steps);
    else
      s->run_dopri(max_steps);
  });
}

int hh_stepper_state(hh_stepper *s, float *x, float *v, double *t, int64_t counts[4]) {
  return guarded([&]() {
    if (!need(s)->initialized) throw HHError(HH_ERR_STATE, "initialize has not run");
    HIPCHK(hipStreamSynchronize(s->st));
    const size_t b = 12 * (size_t)s->n;
    if (x) HIPCHK(hipMemcpy(x, s->X.p, b, hipMemcpyDeviceToHost));
    if (v) HIPCHK(hipMemcpy(v, s->V.p, b, hipMemcpyDeviceToHost));
    if (t) *t = s->t;
    if (counts) {
      counts[0] = s->evals;
      counts[1] = s->count;
      counts[2] = s->rejected;
      counts[3] = s->full_launches;
    }
  });
}

}  // extern "C"
