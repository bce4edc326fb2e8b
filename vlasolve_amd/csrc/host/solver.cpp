// Solver surface: set-up on the host (feet, delay lines, renormalisation),
// every step on the GPU through the vls_* C ABI, and the reference's run loop,
// statistics, snapshots and restarts around it.
#pragma clang fp contract(off)

#include <sys/times.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "vls_host.hpp"

namespace vlh {

namespace {

void check(int rc) {
  if (rc) throw Failure(vls_last_error(), rc);
}

// ---------------------------------------------------------------- binary I/O
// myIO::writeTag/checkTag (src/myIO.hxx:48-87): int length, then the bytes.
void writeTag(FILE *f, const std::string &s) {
  int n = (int)s.size();
  fwrite(&n, sizeof(int), 1, f);
  fwrite(s.data(), 1, n, f);
}

struct Reader {
  FILE *f;
  bool swap = false;
  std::string name;
  void raw(void *p, size_t sz, size_t n) {
    if (fread(p, sz, n, f) != n) throw Failure("ERROR reading file " + name + ": truncated", VLS_ERR_IO);
    if (!swap || sz == 1) return;
    unsigned char *c = (unsigned char *)p;
    for (size_t i = 0; i < n; i++)
      for (size_t a = 0, b = sz - 1; a < b; a++, b--) std::swap(c[i * sz + a], c[i * sz + b]);
  }
  template <class T>
  T get() {
    T v;
    raw(&v, sizeof(T), 1);
    return v;
  }
  void tag(const std::string &ref) {
    int n = 0;
    raw(&n, sizeof(int), 1);
    if (n != (int)ref.size())
      throw Failure("\nERROR reading file: tag '" + ref + "' do not match (size)!!!", VLS_ERR_IO);
    std::string s(n, '\0');
    raw(&s[0], 1, n);
    if (s != ref) throw Failure("\nERROR reding file: tags '" + ref + "' do not match !!!", VLS_ERR_IO);
  }
  void skip(size_t n) {
    std::vector<char> b(n);
    raw(b.data(), 1, n);
  }
};

const char *kTagSolver = "VLASOV_SOLVER v0.13";
const char *kTagHandler = "GRID_HANDLER v0.11";
const char *kTagBase = "GRID_BASE v0.12";
const char *kTag21 = "GRID_21 v0.11";
const char *kTagSolver21 = "VLASOV_SOLVER21 v0.12";
const char *kTagKernel = "KERNEL_CELLS v0.10";
const char *kTagSpline = "VLASOV_SOLVER21_SPLINE v0.11";

// gridBase::paramT::write (src/grid_base.hxx:123-153), including its counts:
// Pmax is written with the v dimension count and Umin with the r one, so the
// file holds Pmin, Pmax, Umin[0], Umin[0], Umax[0], Umax[1].
void writeGridParams(FILE *f, const GridParams &p, int nFields) {
  double d[6] = {p.Pmin, p.Pmax, p.Umin[0], p.Umin[0], p.Umax[0], p.Umax[1]};
  fwrite(d, sizeof(double), 6, f);
  int iv[] = {p.Pres, p.Ures[0], p.Ures[1], p.Pscale, p.Uscale[0], p.Uscale[1],
              0, 0, 0, 0, 0, 0,  // low/high margins
              nFields, 0, 1, 2, 0, 1, 2};
  fwrite(iv, sizeof(int), sizeof(iv) / sizeof(int), f);
  long minStorage = 0;
  fwrite(&minStorage, sizeof(long), 1, f);
  std::vector<char> spare(256 * 8 - sizeof(int) * (3 + 3) - sizeof(long), 0);
  fwrite(spare.data(), 1, spare.size(), f);
}

GridParams readGridParams(Reader &rd, double uminJ) {
  GridParams p;
  double d[6];
  rd.raw(d, sizeof(double), 6);
  p.Pmin = d[0];
  p.Pmax = d[1];
  p.Umin[0] = d[3];
  p.Umax[0] = d[4];
  p.Umax[1] = d[5];
  p.Umin[1] = uminJ;  // not stored in the file (see writeGridParams)
  int iv[19];
  rd.raw(iv, sizeof(int), 19);
  p.Pres = iv[0];
  p.Ures[0] = iv[1];
  p.Ures[1] = iv[2];
  p.Pscale = iv[3];
  p.Uscale[0] = iv[4];
  p.Uscale[1] = iv[5];
  rd.get<long>();
  rd.skip(256 * 8 - sizeof(int) * (3 + 3) - sizeof(long));
  return p;
}

}  // namespace

// ---------------------------------------------------------------- feet
// vlasovSolver21::advanceParticleDrift (src/vlasovSolver21.hxx:328-408).
// pow(x,2) of the reference is its C++98 pow(double,int): the exact square.
bool advanceDrift(double R, double U, double J, double dt, double R0, double &oR, double &oU) {
  double H0, newU, newR, tlim, tR0, sgn, deltaT, t;
  H0 = 0.5 * (U * U + (J * J) / (R * R));
  sgn = (U <= 0) ? -1 : 1;
  deltaT = dt;
  newR = R;
  newU = U;
  if (sgn < 0) {
    tlim = 10 * dt;
    tR0 = 10 * dt;
  } else {
    tlim = std::sqrt(2 * R * R * H0 - J * J) / (2 * H0);
    if (R0 > 0) {
      double A = 2 * H0;
      double B = -2 * sgn * std::sqrt(2 * R * R * H0 - J * J);
      double C = (R * R - R0 * R0);
      double D = B * B - 4 * A * C;
      if (D < 0) tR0 = 10 * dt;
      else tR0 = (-B - std::sqrt(D)) / (2 * A);
      if (tR0 < 0) {
        tR0 = (-B + std::sqrt(D)) / (2 * A);
        if (tR0 < 0) tR0 = 10 * dt;
      }
    } else tR0 = 10 * dt;
  }
  if ((tR0 < deltaT) && (tR0 < tlim)) {  // reflection at R0
    t = std::fabs(newR * newU) - 2 * sgn * H0 * tR0;
    newR = std::sqrt(t * t + J * J) * std::pow(2 * H0, -0.5);
    newU = (2 * H0 - (J * J) / (newR * newR));
    newU = -std::sqrt(newU);
    sgn = -1;
    H0 = 0.5 * (newU * newU + (J * J) / (newR * newR));
    deltaT -= tR0;
    tlim = 10 * dt;
    tR0 = 10 * dt;
  }
  if (tlim < deltaT) {  // pericentre passage
    t = std::fabs(newR * newU) - 2 * sgn * H0 * tlim;
    newR = std::sqrt(t * t + J * J) * std::pow(2 * H0, -0.5);
    newU = 0;
    H0 = 0.5 * (J * J) / (newR * newR);
    sgn = -1;
    deltaT -= tlim;
    tR0 -= tlim;
    tlim = 10 * dt;
    if (tR0 < deltaT) return false;  // the reference aborts here (:389-399)
  }
  t = std::fabs(newR * newU) - 2 * sgn * H0 * deltaT;
  newR = std::sqrt(t * t + J * J) * std::pow(2 * H0, -0.5);
  newU = (2 * H0 - (J * J) / (newR * newR));
  if (newU <= 0) newU = 0;
  else newU = sgn * std::sqrt(newU);
  oR = newR;
  oU = newU;
  return true;
}

// ---------------------------------------------------------------- solver
Solver21::~Solver21() {
  if (ctx_) vls_destroy(ctx_);
}

bool Solver21::checkTime() const {
  // vlasovSolver::checkTime (src/vlasovSolver.hxx:218-221)
  return ((sp_.Tmax <= 0) || ((time_ < sp_.Tmax) && (std::fabs(time_ - sp_.Tmax) > sp_.dt / 100)));
}

void Solver21::init(const Params &prm, const std::string &reload, bool device) {
  if (!reload.empty()) {
    readRestart(reload, prm, device);
    return;
  }
  // vlasovSolver::init (src/vlasovSolver.hxx:61-102)
  initType_ = prm.get<std::string>("type", "init", "UDF");
  std::unique_ptr<InitGen> init = makeInitGen(initType_);
  sp_ = init->defaultSolver();
  sp_.G = prm.get<double>("G", "solver", sp_.G);
  sp_.T0 = prm.get<double>("T0", "solver", sp_.T0);
  sp_.Tmax = prm.get<double>("Tmax", "solver", sp_.Tmax);
  sp_.dt = prm.get<double>("dt", "solver", sp_.dt);
  sp_.nSpecies = prm.get<int>("nSpecies", "solver", sp_.nSpecies);
  sp_.statsEvery = prm.get<long>("statsEvery", "solver", 25);
  sp_.snapshotEvery = prm.get<long>("snapshotEvery", "solver", 500);
  sp_.restartEvery = prm.get<long>("restartEvery", "solver", 5000);
  noLinking_ = prm.get<bool>("noLinking", "solver", false);
  if (sp_.nSpecies < 1 || sp_.nSpecies > init->maxSpecies())
    throw Failure("ERROR: solver.nSpecies " + std::to_string(sp_.nSpecies) + " is not supported by " +
                  initType_ + " initial conditions (at most " + std::to_string(init->maxSpecies()) + ")");
  time_ = sp_.T0;
  init->setup(prm, sp_);
  // gridHandler overrides (src/grid_handler.hxx:221-234)
  GridParams gp = init->defaultGrid();
  gp.Pmin = prm.get<double>("Pmin", "grid", gp.Pmin, 0);
  gp.Pmax = prm.get<double>("Pmax", "grid", gp.Pmax, 0);
  gp.Pres = prm.get<int>("Pres", "grid", gp.Pres, 0);
  gp.Pscale = scaleFromString(prm.get<std::string>("Pscale", "grid", scaleToString(gp.Pscale), 0));
  for (int i = 0; i < 2; i++) {
    gp.Umin[i] = prm.get<double>("Umin", "grid", gp.Umin[i], i);
    gp.Umax[i] = prm.get<double>("Umax", "grid", gp.Umax[i], i);
    gp.Ures[i] = prm.get<int>("Ures", "grid", gp.Ures[i], i);
    gp.Uscale[i] = scaleFromString(prm.get<std::string>("Uscale", "grid", scaleToString(gp.Uscale[i]), i));
  }
  grid_.init(gp);
  f_.assign(grid_.nVal * sp_.nSpecies, 0.0);
  if (verbose) { printf("Generating initial conditions (%s[%d]) ... ", initType_.c_str(), sp_.nSpecies); fflush(stdout); }
  init->generate(grid_, f_.data(), sp_.nSpecies);
  if (verbose) printf("done.\n");
  // vlasovSolver21::setup (src/vlasovSolver21.hxx:533-548)
  std::string kt = prm.get<std::string>("kernelType", "solver", "delayed");
  if (kt == "delayed") kernelType_ = DELAYED;
  else if (kt == "reflective") kernelType_ = REFLECTIVE;
  else throw Failure("ERROR: unknown kernel type value: '" + kt + "'");
  kernelBufferSize_ = prm.get<int>("kernelBufferSize", "solver", 0);
  kernelMass_ = 0;
  setupFeet(&init);
  // vlasovSolver::init renormalizes each species to its mass (src/vlasovSolver.hxx:94-95)
  for (int s = 0; s < sp_.nSpecies; s++) renormalize(init->mass(s), s);
  initGen_ = std::move(init);
  if (device) upload();
}

void Solver21::setupFeet(std::unique_ptr<InitGen> *init) {
  const Grid21 &g = grid_;
  if (kernelBufferSize_ < 0 || kernelBufferSize_ >= g.nR) throw Failure("ERROR: kernelBufferSize out of range");
  kernelRadius_ = g.R_Val[kernelBufferSize_];
  if (verbose) printf("kernel radius: %e\n", kernelRadius_);
  // generateDisplacement (src/vlasovSolver21.hxx:102-143); with linking the
  // feet span a whole dt (setup() passes 2 dt, :443-444)
  const double dtFeet = noLinking_ ? sp_.dt : sp_.dt * 2;
  const double dth = 0.5 * dtFeet;
  const double R0 = (kernelType_ == REFLECTIVE) ? kernelRadius_ : -1;
  driftR_.resize(g.nVal);
  driftU_.resize(g.nVal);
  driftOut_.assign(g.nVal, 0);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int j = 0; j < g.nJ; j++)
    for (int u = 0; u < g.nU; u++)
      for (int r = 0; r < g.nR; r++) {
        int64_t i = g.index(j, u, r);
        double nr = 0, nu = 0;
        if (!advanceDrift(g.R_Val[r], g.U_Val[u], g.J_Val[j], dth, R0, nr, nu)) bad = 1;
        driftR_[i] = nr;
        driftU_[i] = nu;
        driftOut_[i] = (nr >= g.R_Val.back()) || (nr <= g.R_Val.front()) || (nu >= g.U_Val.back()) ||
                       (nu <= g.U_Val.front());
      }
  if (bad) throw Failure("ERROR: a foot crosses both the pericentre and the kernel radius in one step");
  if (verbose) printf("kernel mass : %e\n", kernelMass_);
  chains_ = Chains();
  if (kernelType_ != DELAYED || !init) {
    if (kernelType_ != DELAYED)
      for (int64_t i = 0; i < g.nVal; i++)
        if (driftR_[i] < kernelRadius_) throw Failure("ERROR: reflective kernel: a foot falls inside the kernel radius");
    return;
  }
  // setupBoundaryConditions (src/vlasovSolver21.hxx:265-326): one delay line
  // per point whose foot falls inside the kernel radius, holding the initial
  // distribution at the positions the orbit had at earlier multiples of dth
  for (int j = 0; j < g.nJ; j++)
    for (int u = 0; u < g.nU; u++)
      for (int r = 0; r < g.nR; r++) {
        int64_t i = g.index(j, u, r);
        if (!(driftR_[i] < kernelRadius_)) continue;
        const double J = g.J_Val[j];
        chains_.point.push_back(i);
        chains_.off.push_back((int64_t)chains_.cellR.size());
        double deltaT = dth, nr, nu;
        if (!advanceDrift(g.R_Val[r], g.U_Val[u], J, deltaT, -1, nr, nu)) throw Failure("ERROR: chain set-up failed");
        int32_t n = 0;
        while (nr < kernelRadius_) {
          chains_.cellR.push_back(nr);
          chains_.cellU.push_back(nu);
          for (int s = 0; s < sp_.nSpecies; s++) chains_.val.push_back((*init)->valueAt(nr, nu, J, s));
          n++;
          deltaT += dth;
          if (!advanceDrift(g.R_Val[r], g.U_Val[u], J, deltaT, -1, nr, nu)) throw Failure("ERROR: chain set-up failed");
        }
        // the back cell starts at 0 for every species (the reference makes it
        // with one value, dummyCell21(R,U,1), src/vlasovSolver21.hxx:322, so
        // that with several species its species >= 1 fall outside the cell)
        chains_.cellR.push_back(nr);
        chains_.cellU.push_back(nu);
        for (int s = 0; s < sp_.nSpecies; s++) chains_.val.push_back(0.0);
        chains_.len.push_back(n + 1);
        chains_.head.push_back(0);
      }
}

double Solver21::quadratureMass(const double *f, int nS, int sp) const {
  // quadrature21::integrate (src/quadrature21.hxx:57-85) summed in the
  // reference's single-thread order
  const Grid21 &g = grid_;
  double mass = 0;
  for (int j = 0; j < g.nJ; j++) {
    double Ct1 = 8 * M_PI * M_PI * g.J_Val[j] * g.dJ[j];
    for (int u = 0; u < g.nU; u++)
      for (int r = 0; r < g.nR; r++) {
        double dU = g.dU[u];
        if (dU <= 0) continue;
        double dR = g.dR[r];
        if (dR <= 0) continue;
        const double *c = &f[g.index(j, u, r) * nS + sp];
        double avg = ((*c) + c[nS] + c[(int64_t)g.nR * nS] + c[(int64_t)(g.nR + 1) * nS]) * 0.25;
        mass += Ct1 * dU * dR * avg;
      }
  }
  return mass;
}

void Solver21::renormalize(double norm, int sp) {
  // vlasovSolver21::renormalize(norm, which = sp) (src/vlasovSolver21.hxx:468-524)
  // with the mass of species sp
  if (norm <= 0) return;
  const int nS = sp_.nSpecies;
  double mass = quadratureMass(f_.data(), nS, sp);
  mass = mass + kernelMass_;
  double fac = norm / mass;
  if (verbose) printf("Renormalizing mass : M[%d]=%g -> %g.\n", sp, mass, norm);
  for (size_t i = sp; i < f_.size(); i += nS) f_[i] *= fac;
  for (size_t i = sp; i < chains_.val.size(); i += nS) chains_.val[i] *= fac;
}

void Solver21::initDerivative(const std::string &param, int s, std::vector<double> &df,
                              std::vector<double> &dvals) const {
  if (!initGen_) throw Failure("ERROR: no initial conditions to differentiate (a reloaded run)");
  const int nS = sp_.nSpecies;
  if (s < 0 || s >= nS) throw Failure("ERROR: species index out of range");
  const Grid21 &g = grid_;
  // the generator's values (before the renormalisation) and their derivatives
  std::vector<double> raw(g.nVal), draw(g.nVal);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int j = 0; j < g.nJ; j++)
    for (int u = 0; u < g.nU; u++)
      for (int r = 0; r < g.nR; r++) {
        const int64_t i = g.index(j, u, r);
        if (!initGen_->dValueAt(g.R_Val[r], g.U_Val[u], g.J_Val[j], s, param, raw[i], draw[i]))
          bad = 1;
      }
  if (bad)
    throw Failure("ERROR: " + initType_ + " initial conditions have no derivative in '" + param + "'");
  // f0 = raw * fac, fac = M / (mass(raw) + kernelMass): the mass is linear in f
  const double Q = quadratureMass(raw.data(), 1, 0) + kernelMass_;
  const double dQ = quadratureMass(draw.data(), 1, 0);
  const double M = initGen_->mass(s), dM = param == "M" ? 1.0 : 0.0;
  const double fac = M > 0 ? M / Q : 1.0, dfac = M > 0 ? dM / Q - M * dQ / (Q * Q) : 0.0;
  df.assign(f_.size(), 0.0);
  for (int64_t i = 0; i < g.nVal; i++) df[i * nS + s] = draw[i] * fac + raw[i] * dfac;
  // the delay-line cells hold the generator's values at their positions,
  // scaled by the same factor; each line's back cell starts at 0
  const Chains &c = chains_;
  dvals.assign(c.val.size(), 0.0);
  for (size_t k = 0; k < c.point.size(); k++) {
    const int j = (int)(c.point[k] / ((int64_t)g.nU * g.nR));
    for (int32_t q = 0; q + 1 < c.len[k]; q++) {
      const int64_t cell = c.off[k] + q;
      double v = 0, dv = 0;
      initGen_->dValueAt(c.cellR[cell], c.cellU[cell], g.J_Val[j], s, param, v, dv);
      dvals[cell * nS + s] = dv * fac + v * dfac;
    }
  }
}

vls_solver_params Solver21::deviceParams() const {
  vls_solver_params sp{};
  sp.G = sp_.G;
  sp.dt = sp_.dt;
  sp.no_linking = noLinking_ ? 1 : 0;
  sp.kernel_type = kernelType_;
  sp.n_species = sp_.nSpecies;
  return sp;
}

void Solver21::upload() {
  if (ctx_) {
    vls_destroy(ctx_);
    ctx_ = nullptr;
  }
  vls_geometry geo = grid_.geometry();
  vls_solver_params sp = deviceParams();
  check(vls_create(&geo, &sp, 1, &ctx_));
  uploadTo(ctx_, 0);
}

void Solver21::uploadTo(vls_ctx *ctx, int run) const {
  const Grid21 &g = grid_;
  check(vls_upload_f(ctx, run, f_.data()));
  std::vector<double> xr(g.nVal, 0.0), xu(g.nVal, 0.0);
  std::vector<uint8_t> flag(g.nVal);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < g.nVal; i++) {
    if (driftR_[i] < kernelRadius_) flag[i] = VLS_FOOT_CHAIN;
    else if (driftOut_[i]) flag[i] = VLS_FOOT_OUT;
    else {
      flag[i] = VLS_FOOT_EVAL;
      xr[i] = g.splineCoord(driftR_[i], 0);
      xu[i] = g.splineCoord(driftU_[i], 1);
    }
  }
  check(vls_set_drift(ctx, run, xr.data(), xu.data(), flag.data()));
  const int64_t nc = (int64_t)chains_.point.size();
  std::vector<double> bxr(nc), bxu(nc), ringVals(chains_.val.size());
  std::vector<uint8_t> bout(nc);
  for (int64_t c = 0; c < nc; c++) {
    int64_t b = chains_.off[c] + chains_.len[c] - 1;
    double U = chains_.cellU[b];
    bout[c] = (U <= g.U_Val.front()) || (U >= g.U_Val.back());
    bxr[c] = bout[c] ? 0.0 : g.splineCoord(chains_.cellR[b], 0);
    bxu[c] = bout[c] ? 0.0 : g.splineCoord(U, 1);
    // ring storage: slot (head + k) % len holds list position k
    const int nS = sp_.nSpecies;
    for (int32_t k = 0; k < chains_.len[c]; k++)
      for (int s = 0; s < nS; s++)
        ringVals[(chains_.off[c] + (chains_.head[c] + k) % chains_.len[c]) * nS + s] =
            chains_.val[(chains_.off[c] + k) * nS + s];
  }
  check(vls_set_chains(ctx, run, nc, chains_.point.data(), chains_.len.data(), chains_.head.data(),
                       bxr.data(), bxu.data(), bout.data(), ringVals.data()));
}

void Solver21::makeSteps(int n) {
  check(vls_step(ctx_, n));
  for (int k = 0; k < n; k++) {
    time_ += sp_.dt;
    stepCount_++;
  }
}

void Solver21::stats(double out[7]) { check(vls_stats(ctx_, 0, out)); }
void Solver21::statsTotal(double out[7]) { check(vls_stats_total(ctx_, 0, nullptr, out)); }
void Solver21::downloadF(double *f) { check(vls_download_f(ctx_, 0, f)); }
void Solver21::moments(double *dM, double *accel, double *mTot) {
  check(vls_moments(ctx_, 0, dM, accel, mTot));
}

void Solver21::statistics(bool reset) {
  // vlasovSolver21::statistics (src/vlasovSolver21.hxx:644-711): with several
  // species one file per species (the same line in each, as the reference
  // computes it), then statistics.txt with the total line
  double line[7], tot[7];
  check(vls_stats_total(ctx_, 0, line, tot));
  auto put = [&](const char *fname, const double *s) {
    FILE *f = fopen(fname, reset ? "w" : "a");
    if (!f) throw Failure(std::string("ERROR: cannot open ") + fname, VLS_ERR_IO);
    if (reset) fprintf(f, "# Time Step Mass E Ek Ep S L1 L2\n");
    fprintf(f, "%.10g %ld %10.10g %10.10g %10.10g %10.10g %10.10g %10.10g %10.10g\n", time_, stepCount_,
            s[0], s[1], s[2], s[3], s[4], s[5], s[6]);
    fclose(f);
  };
  if (sp_.nSpecies > 1) {
    for (int s = 0; s < sp_.nSpecies; s++) {
      char fname[64];
      snprintf(fname, sizeof fname, "statistics_S%d.txt", s);
      put(fname, line);
    }
    put("statistics.txt", tot);
  } else {
    put("statistics.txt", line);
  }
}

void Solver21::snapshot() {
  // vlasovSolver21::snapshot + NDfieldFromData (src/vlasovSolver21.hxx:585-598,717-773)
  // (with several species a first, fastest dimension of nSpecies: :719-727)
  downloadF(f_.data());
  const GridParams &p = grid_.p;
  const int sd = sp_.nSpecies > 1 ? 1 : 0;
  int dims[4] = {sp_.nSpecies, p.Pres + 1, p.Ures[0] + 1, p.Ures[1]};
  double x0[4] = {0, p.Pmin, p.Umin[0], p.Umin[1]};
  double delta[4] = {(double)sp_.nSpecies, p.Pmax - p.Pmin, p.Umax[0] - p.Umin[0], p.Umax[1] - p.Umin[1]};
  char comment[80];
  snprintf(comment, sizeof comment, "%d %d %d %d %d %d", p.Pscale, 1, p.Uscale[0], 1, p.Uscale[1], 0);
  char fname[1024];
  snprintf(fname, sizeof fname, "snapshot_%4.4f.ND", time_);
  saveNDfield(fname, dims + 1 - sd, 3 + sd, x0 + 1 - sd, delta + 1 - sd, f_.data(), comment);
}

void Solver21::write(const std::string &fnameNoExt) {
  // vlasovSolver::write chain (src/vlasovSolver.hxx:248-287 ->
  // gridHandler::write src/grid_handler.hxx:325-347 -> gridBase/grid21 ->
  // vlasovSolver21::write src/vlasovSolver21.hxx:568-583 -> spline tag)
  downloadF(f_.data());
  const int64_t nc = (int64_t)chains_.point.size();
  std::vector<int32_t> head(nc);
  std::vector<double> ring(chains_.val.size());  // cells x species
  if (nc) check(vls_get_chains(ctx_, 0, head.data(), ring.data()));
  const int nSp = sp_.nSpecies;
  std::string fname = fnameNoExt + ".vls";
  FILE *f = fopen(fname.c_str(), "w");
  if (!f) throw Failure("ERROR: opening file '" + fname + "' for writing.", VLS_ERR_IO);
  if (verbose) { printf("\nDumping solver state (%s) ... ", fname.c_str()); fflush(stdout); }
  int one = 1;
  fwrite(&one, sizeof(int), 1, f);
  writeTag(f, kTagSolver);
  fwrite(&sp_.G, sizeof(double), 1, f);
  fwrite(&sp_.T0, sizeof(double), 1, f);
  fwrite(&sp_.Tmax, sizeof(double), 1, f);
  fwrite(&sp_.dt, sizeof(double), 1, f);
  fwrite(&sp_.nSpecies, sizeof(int), 1, f);
  fwrite(&sp_.snapshotEvery, sizeof(long), 1, f);
  fwrite(&sp_.restartEvery, sizeof(long), 1, f);
  fwrite(&sp_.statsEvery, sizeof(long), 1, f);
  fwrite(&time_, sizeof(double), 1, f);
  fwrite(&sp_.dt, sizeof(double), 1, f);
  fwrite(&stepCount_, sizeof(long), 1, f);
  int nl = noLinking_ ? 1 : 0;
  fwrite(&nl, sizeof(int), 1, f);
  // grid handler
  writeTag(f, kTagHandler);
  fwrite(&one, sizeof(int), 1, f);
  int dimsPUJ[3] = {1, 2, 1};
  fwrite(dimsPUJ, sizeof(int), 3, f);
  long l = sizeof(double);
  fwrite(&l, sizeof(long), 1, f);
  l = 1;
  fwrite(&l, sizeof(long), 1, f);  // nGrids
  l = 0;
  fwrite(&l, sizeof(long), 1, f);  // rank
  writeGridParams(f, grid_.p, sp_.nSpecies);
  writeTag(f, kTagBase);
  writeGridParams(f, grid_.p, sp_.nSpecies);
  fwrite(f_.data(), sizeof(double), f_.size(), f);
  writeTag(f, kTag21);
  std::vector<char> spare(256 * 8, 0);
  fwrite(spare.data(), 1, spare.size(), f);
  fwrite(spare.data(), 1, spare.size(), f);  // gridHandler spare
  fwrite(spare.data(), 1, spare.size(), f);  // vlasovSolver spare
  // vlasovSolver21
  writeTag(f, kTagSolver21);
  int kt = kernelType_;
  fwrite(&kt, sizeof(int), 1, f);
  fwrite(&kernelMass_, sizeof(double), 1, f);
  fwrite(&kernelBufferSize_, sizeof(int), 1, f);
  fwrite(spare.data(), 1, spare.size(), f);
  if (kernelType_ == DELAYED) {
    // sphericalKernel::write (src/sphericalKernel.hxx:88-106): per slice, the
    // chains keyed by grid index, cells front first
    writeTag(f, kTagKernel);
    long nS = grid_.nJ;
    fwrite(&nS, sizeof(long), 1, f);
    const int64_t per = (int64_t)grid_.nU * grid_.nR;
    int64_t c = 0;
    for (long j = 0; j < nS; j++) {
      int64_t c0 = c;
      while (c < nc && chains_.point[c] / per == j) c++;
      long cnt = (long)(c - c0);
      fwrite(&cnt, sizeof(long), 1, f);
      for (int64_t k = c0; k < c; k++) {
        long key = (long)chains_.point[k];
        fwrite(&key, sizeof(long), 1, f);
        long L = chains_.len[k];
        fwrite(&L, sizeof(long), 1, f);
        for (long m = 0; m < L; m++) {
          int64_t o = chains_.off[k];
          fwrite(&chains_.cellR[o + m], sizeof(double), 1, f);
          fwrite(&chains_.cellU[o + m], sizeof(double), 1, f);
          fwrite(&nSp, sizeof(int), 1, f);
          fwrite(&ring[(o + (head[k] + m) % L) * nSp], sizeof(double), nSp, f);
        }
      }
    }
  }
  writeTag(f, kTagSpline);
  fwrite(spare.data(), 1, spare.size(), f);
  fclose(f);
  if (verbose) printf("done.\n");
}

void Solver21::readRestart(const std::string &fnameNoExt, const Params &prm, bool device) {
  // vlasovSolver::init(params, fname) and the read chain
  // (src/vlasovSolver.hxx:104-132, 289-314; vlasovSolver21.hxx:550-566)
  std::string fname = fnameNoExt + ".vls";
  FILE *f = fopen(fname.c_str(), "r");
  if (!f) throw Failure("ERROR: File " + fname + " does not exist.", VLS_ERR_IO);
  Reader rd{f, false, fname};
  try {
    int e = rd.get<int>();
    rd.swap = (e != 1);
    rd.tag(kTagSolver);
    sp_.G = rd.get<double>();
    sp_.T0 = rd.get<double>();
    sp_.Tmax = rd.get<double>();
    sp_.dt = rd.get<double>();
    sp_.nSpecies = rd.get<int>();
    sp_.snapshotEvery = rd.get<long>();
    sp_.restartEvery = rd.get<long>();
    sp_.statsEvery = rd.get<long>();
    time_ = rd.get<double>();
    sp_.dt = rd.get<double>();
    stepCount_ = rd.get<long>();
    noLinking_ = rd.get<int>() != 0;
    rd.tag(kTagHandler);
    rd.get<int>();
    int dimsPUJ[3];
    rd.raw(dimsPUJ, sizeof(int), 3);
    if (dimsPUJ[0] != 1 || dimsPUJ[1] != 2 || dimsPUJ[2] != 1)
      throw Failure("ERROR: container and file have different dimensions.", VLS_ERR_IO);
    if (rd.get<long>() != (long)sizeof(double)) throw Failure("ERROR: container and file have different data types.", VLS_ERR_IO);
    long nGrids = rd.get<long>();
    if (nGrids != 1) throw Failure("ERROR: this file can only be loaded on a " + std::to_string(nGrids) + " nodes MPI run.", VLS_ERR_IO);
    if (rd.get<long>() != 0) throw Failure("ERROR: wrong rank", VLS_ERR_IO);
    // j-axis minimum is absent from the file; take it from the parameters
    double uminJ = prm.get<double>("Umin", "grid", 0.0, 1);
    GridParams full = readGridParams(rd, uminJ);
    rd.tag(kTagBase);
    GridParams gp = readGridParams(rd, uminJ);
    (void)full;
    grid_.init(gp);
    if (sp_.nSpecies < 1 || sp_.nSpecies > 5) throw Failure("ERROR: restart holds an invalid number of species", VLS_ERR_IO);
    f_.resize(grid_.nVal * sp_.nSpecies);
    rd.raw(f_.data(), sizeof(double), f_.size());
    rd.tag(kTag21);
    rd.skip(256 * 8 * 3);
    rd.tag(kTagSolver21);
    kernelType_ = rd.get<int>();
    kernelMass_ = rd.get<double>();
    // The reference adds kernelMass to the enclosed mass of the kick and of
    // the potential energy (vlasovSolver21_spline.hxx:738-742,
    // vlasovSolver21.hxx:632) and always writes 0 (vlasovSolver21.hxx:542-543);
    // the device step has no such term, so a file that carries another value
    // is refused rather than stepped differently.
    if (kernelMass_ != 0.0)
      throw Failure("ERROR: restart holds a nonzero kernelMass, which this solver does not model",
                    VLS_ERR_IO);
    kernelBufferSize_ = rd.get<int>();
    rd.skip(256 * 8);
    setupFeet(nullptr);  // feet are recomputed; chains come from the file
    chains_ = Chains();
    if (kernelType_ == DELAYED) {
      rd.tag(kTagKernel);
      long nS = rd.get<long>();
      struct C { long key; std::vector<double> R, U, v; };
      std::vector<C> all;
      for (long j = 0; j < nS; j++) {
        long cnt = rd.get<long>();
        for (long k = 0; k < cnt; k++) {
          C c;
          c.key = rd.get<long>();
          long L = rd.get<long>();
          for (long m = 0; m < L; m++) {
            c.R.push_back(rd.get<double>());
            c.U.push_back(rd.get<double>());
            int nv = rd.get<int>();
            std::vector<double> v(nv);
            rd.raw(v.data(), sizeof(double), nv);
            // one value per species; a cell the reference made with fewer
            // (its back cell, dummyCell21(R,U,1)) holds 0 for the others
            for (int s = 0; s < sp_.nSpecies; s++) c.v.push_back(s < nv ? v[s] : 0.0);
          }
          all.push_back(c);
        }
      }
      std::sort(all.begin(), all.end(), [](const C &a, const C &b) { return a.key < b.key; });
      for (auto &c : all) {
        if (c.key < 0 || c.key >= grid_.nVal || !(driftR_[c.key] < kernelRadius_))
          throw Failure("ERROR: restart chain does not match the recomputed feet", VLS_ERR_IO);
        chains_.point.push_back(c.key);
        chains_.off.push_back((int64_t)chains_.cellR.size());
        chains_.len.push_back((int32_t)c.R.size());
        chains_.head.push_back(0);
        chains_.cellR.insert(chains_.cellR.end(), c.R.begin(), c.R.end());
        chains_.cellU.insert(chains_.cellU.end(), c.U.begin(), c.U.end());
        chains_.val.insert(chains_.val.end(), c.v.begin(), c.v.end());
      }
    }
    rd.tag(kTagSpline);
    rd.skip(256 * 8);
  } catch (...) {
    fclose(f);
    throw;
  }
  fclose(f);
  sp_.Tmax = prm.get<double>("Tmax", "solver", sp_.Tmax);
  sp_.statsEvery = prm.get<long>("statsEvery", "solver", sp_.statsEvery);
  sp_.snapshotEvery = prm.get<long>("snapshotEvery", "solver", sp_.snapshotEvery);
  sp_.restartEvery = prm.get<long>("restartEvery", "solver", sp_.restartEvery);
  if (device) upload();
}

void Solver21::run() {
  // vlasovSolver::run (src/vlasovSolver.hxx:134-190)
  if (!checkTime()) {
    printf("WARNING: nothing to do, T0 = %.3f >= Tmax = %.3f.\n", time_, sp_.Tmax);
    return;
  }
  printf("Using %s kernel with M0=%e.\n", kernelType_ == DELAYED ? "delayed" : "reflective", kernelMass_);
  printf("Starting at T=%.5f, dt=%g, tmax = %g\n", time_, sp_.dt, sp_.Tmax);
  fflush(stdout);
  auto t0 = std::chrono::steady_clock::now();
  long stepCount0 = stepCount_;
  do {
    if (checkTimer(sp_.restartEvery)) {
      char fname[255];
      snprintf(fname, sizeof fname, "restart_%.4f", time_);
      write(fname);
    }
    if (checkTimer(sp_.snapshotEvery)) snapshot();
    if (checkTimer(sp_.statsEvery)) {
      statistics(stepCount_ == 0);
      if (verbose) { printf("\rT=%.4f -> %.4f   ", time_, time_ + sp_.dt); fflush(stdout); }
    }
    makeSteps(1);
  } while (checkTime());
  check(vls_synchronize(ctx_));
  printf("\n");
  double T = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  printf("Computed %ld timesteps in %gs. (%3.3gs / step)", stepCount_, T, T / (stepCount_ - stepCount0));
  // afterRunning (src/vlasovSolver.hxx:235-246)
  if (checkTimer(sp_.snapshotEvery)) snapshot();
  char fname[255];
  snprintf(fname, sizeof fname, "final_%.4f", time_);
  write(fname);
}

// ---------------------------------------------------------------- NDfield
void saveNDfield(const std::string &fname, const int *dims, int ndims, const double *x0,
                 const double *delta, const double *data, const std::string &comment) {
  FILE *f = fopen(fname.c_str(), "w");
  if (!f) throw Failure("ERROR: cannot write " + fname, VLS_ERR_IO);
  if (verbose) { printf("Saving [%d,%d,%d] DOUBLE array to file %s ...", dims[0], dims[1], dims[2], fname.c_str()); fflush(stdout); }
  const int MAXD = 20;
  int i = 16;
  char tag[16] = {0};
  strcpy(tag, "NDFIELD");
  fwrite(&i, sizeof(int), 1, f);
  fwrite(tag, 1, 16, f);
  fwrite(&i, sizeof(int), 1, f);
  i = sizeof(int) * (MAXD + 3) + sizeof(double) * (2 * MAXD) + (160 + 80);
  fwrite(&i, sizeof(int), 1, f);
  char cm[80] = {0};
  strncpy(cm, comment.c_str(), 79);
  fwrite(cm, 1, 80, f);
  fwrite(&ndims, sizeof(int), 1, f);
  int dd[MAXD] = {0};
  double x[MAXD] = {0}, dl[MAXD] = {0};
  long nval = 1;
  for (int k = 0; k < ndims; k++) {
    dd[k] = dims[k];
    x[k] = x0[k];
    dl[k] = delta[k];
    nval *= dims[k];
  }
  fwrite(dd, sizeof(int), MAXD, f);
  int fdi = 0, type = 1 << 9;
  fwrite(&fdi, sizeof(int), 1, f);
  fwrite(&type, sizeof(int), 1, f);
  fwrite(x, sizeof(double), MAXD, f);
  fwrite(dl, sizeof(double), MAXD, f);
  char dummy[160] = {0};
  fwrite(dummy, 1, 160, f);
  fwrite(&i, sizeof(int), 1, f);
  int nb = (int)(nval * (long)sizeof(double));
  fwrite(&nb, sizeof(int), 1, f);
  fwrite(data, sizeof(double), nval, f);
  fwrite(&nb, sizeof(int), 1, f);
  fclose(f);
  if (verbose) printf(" done.\n");
}

}  // namespace vlh
