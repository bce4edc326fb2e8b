// C ABI of the solver surface (include/vlasolve_hip.h, layer 2).
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "vls_host.hpp"

struct vlsd_solver {
  vlh::Params prm;
  vlh::Solver21 solver;
};

namespace {
thread_local std::string g_derr;

template <class F>
int guarded(F &&fn) {
  try {
    fn();
    return VLS_OK;
  } catch (const vlh::Failure &e) {
    g_derr = e.what();
    return e.code ? e.code : VLS_ERR_ARG;
  } catch (const std::exception &e) {
    g_derr = e.what();
    return VLS_ERR_ARG;
  }
}
}  // namespace

extern "C" {

const char *vlsd_last_error(void) { return g_derr.c_str(); }

int vlsd_create(int argc, const char *const *argv, vlsd_solver **out) {
  if (!out) return VLS_ERR_ARG;
  *out = nullptr;
  vlsd_solver *s = new vlsd_solver();
  int rc = guarded([&] {
    s->prm = vlh::Params(argc, argv);
    vlh::verbose = s->prm.get<int>("verbose", "default", vlh::verbose);
    std::string reload = s->prm.get<std::string>("reload", "default", "");
    s->solver.init(s->prm, reload);
  });
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return VLS_OK;
}

void vlsd_destroy(vlsd_solver *s) { delete s; }

int vlsd_info(vlsd_solver *s, int64_t dims[3], double *time, int64_t *stepCount) {
  return guarded([&] {
    if (!s) throw vlh::Failure("null solver");
    const vlh::Grid21 &g = s->solver.grid();
    if (dims) {
      dims[0] = g.nR;
      dims[1] = g.nU;
      dims[2] = g.nJ;
    }
    if (time) *time = s->solver.time();
    if (stepCount) *stepCount = s->solver.stepCount();
  });
}

int vlsd_species(vlsd_solver *s) { return s ? s->solver.nSpecies() : 0; }

int vlsd_step(vlsd_solver *s, int nSteps) {
  return guarded([&] {
    if (!s || nSteps < 0) throw vlh::Failure("bad arguments");
    s->solver.makeSteps(nSteps);
  });
}

int vlsd_get_f(vlsd_solver *s, double *f) {
  return guarded([&] {
    if (!s || !f) throw vlh::Failure("bad arguments");
    s->solver.downloadF(f);
  });
}

int vlsd_stats(vlsd_solver *s, double out[7]) {
  return guarded([&] {
    if (!s || !out) throw vlh::Failure("bad arguments");
    s->solver.stats(out);
  });
}

int vlsd_stats_total(vlsd_solver *s, double out[7]) {
  return guarded([&] {
    if (!s || !out) throw vlh::Failure("bad arguments");
    s->solver.statsTotal(out);
  });
}

int vlsd_moments(vlsd_solver *s, double *dM, double *accel, double *mTot) {
  return guarded([&] {
    if (!s) throw vlh::Failure("bad arguments");
    s->solver.moments(dM, accel, mTot);
  });
}

int vlsd_run(vlsd_solver *s) {
  return guarded([&] {
    if (!s) throw vlh::Failure("bad arguments");
    s->solver.run();
  });
}

int vlsd_write_restart(vlsd_solver *s, const char *fname) {
  return guarded([&] {
    if (!s || !fname) throw vlh::Failure("bad arguments");
    s->solver.write(fname);
  });
}

int vlsd_snapshot(vlsd_solver *s) {
  return guarded([&] {
    if (!s) throw vlh::Failure("bad arguments");
    s->solver.snapshot();
  });
}

vls_ctx *vlsd_context(vlsd_solver *s) { return s ? s->solver.context() : nullptr; }

}  // extern "C"

// ---------------------------------------------------------------- host set-up only
struct vlsd_host {
  vlh::Params prm;
  vlh::Solver21 solver;
};

struct vlsd_ensemble {
  std::vector<std::unique_ptr<vlh::Solver21>> runs;
  vls_ctx *ctx = nullptr;
  ~vlsd_ensemble() {
    if (ctx) vls_destroy(ctx);
  }
};

extern "C" {

int vlsd_ensemble_create(int nRuns, const int *argc, const char *const *const *argv,
                         vlsd_ensemble **out) {
  if (!out) return VLS_ERR_ARG;
  *out = nullptr;
  vlsd_ensemble *e = new vlsd_ensemble();
  int rc = guarded([&] {
    if (nRuns < 1 || !argc || !argv) throw vlh::Failure("vlsd_ensemble_create: need runs");
    std::vector<vls_geometry> geo(nRuns);
    std::vector<vls_solver_params> sp(nRuns);
    for (int k = 0; k < nRuns; k++) {
      vlh::Params prm(argc[k], argv[k]);
      vlh::verbose = prm.get<int>("verbose", "default", 0);
      if (!prm.get<std::string>("reload", "default", "").empty())
        throw vlh::Failure("vlsd_ensemble_create: runs start from initial conditions");
      e->runs.emplace_back(new vlh::Solver21());
      e->runs.back()->init(prm, "", false);
      geo[k] = e->runs.back()->grid().geometry();
      sp[k] = e->runs.back()->deviceParams();
    }
    int rc2 = vls_create(geo.data(), sp.data(), nRuns, &e->ctx);
    if (rc2) throw vlh::Failure(vls_last_error(), rc2);
    for (int k = 0; k < nRuns; k++) e->runs[k]->uploadTo(e->ctx, k);
  });
  if (rc) {
    delete e;
    return rc;
  }
  *out = e;
  return VLS_OK;
}

void vlsd_ensemble_destroy(vlsd_ensemble *e) { delete e; }

vls_ctx *vlsd_ensemble_context(vlsd_ensemble *e) { return e ? e->ctx : nullptr; }

int vlsd_host_create(int argc, const char *const *argv, vlsd_host **out) {
  if (!out) return VLS_ERR_ARG;
  *out = nullptr;
  vlsd_host *h = new vlsd_host();
  int rc = guarded([&] {
    h->prm = vlh::Params(argc, argv);
    vlh::verbose = h->prm.get<int>("verbose", "default", vlh::verbose);
    std::string reload = h->prm.get<std::string>("reload", "default", "");
    h->solver.init(h->prm, reload, false);
  });
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return VLS_OK;
}

void vlsd_host_destroy(vlsd_host *h) { delete h; }

int vlsd_host_info(vlsd_host *h, int64_t dims[3], int64_t *nChains, int64_t *nCells,
                   double *kernelRadius) {
  return guarded([&] {
    if (!h) throw vlh::Failure("null handle");
    const vlh::Grid21 &g = h->solver.grid();
    if (dims) {
      dims[0] = g.nR;
      dims[1] = g.nU;
      dims[2] = g.nJ;
    }
    if (nChains) *nChains = (int64_t)h->solver.chains().point.size();
    if (nCells) *nCells = (int64_t)h->solver.chains().cellR.size();
    if (kernelRadius) *kernelRadius = h->solver.kernelRadius();
  });
}

int vlsd_host_species(vlsd_host *h) { return h ? h->solver.nSpecies() : 0; }

int vlsd_host_f(vlsd_host *h, double *f) {
  return guarded([&] {
    if (!h || !f) throw vlh::Failure("null argument");
    const auto &v = h->solver.hostF();
    std::copy(v.begin(), v.end(), f);
  });
}

int vlsd_host_feet(vlsd_host *h, double *driftR, double *driftU, uint8_t *flag) {
  return guarded([&] {
    if (!h) throw vlh::Failure("null handle");
    const vlh::Solver21 &s = h->solver;
    size_t n = s.driftR().size();
    for (size_t i = 0; i < n; i++) {
      if (driftR) driftR[i] = s.driftR()[i];
      if (driftU) driftU[i] = s.driftU()[i];
      if (flag)
        flag[i] = (s.driftR()[i] < s.kernelRadius()) ? VLS_FOOT_CHAIN
                                                       : (s.driftOut()[i] ? VLS_FOOT_OUT : VLS_FOOT_EVAL);
    }
  });
}

int vlsd_host_dparam(vlsd_host *h, const char *param, int species, double *df, double *dvals) {
  return guarded([&] {
    if (!h || !param) throw vlh::Failure("null argument");
    std::vector<double> a, b;
    h->solver.initDerivative(param, species, a, b);
    if (df) std::copy(a.begin(), a.end(), df);
    if (dvals) std::copy(b.begin(), b.end(), dvals);
  });
}

int vlsd_host_chains(vlsd_host *h, int64_t *point, int32_t *len, double *cellR, double *cellU,
                     double *val) {
  return guarded([&] {
    if (!h) throw vlh::Failure("null handle");
    const vlh::Chains &c = h->solver.chains();
    if (point) std::copy(c.point.begin(), c.point.end(), point);
    if (len) std::copy(c.len.begin(), c.len.end(), len);
    if (cellR) std::copy(c.cellR.begin(), c.cellR.end(), cellR);
    if (cellU) std::copy(c.cellU.begin(), c.cellU.end(), cellU);
    if (val) std::copy(c.val.begin(), c.val.end(), val);
  });
}

}  // extern "C"
