// Grid geometry and initial conditions.  The arithmetic of every expression
// follows the reference's (cited per function) so that node values and the
// initial distribution are bit-identical to it on x86-64.
#pragma clang fp contract(off)

#include <cmath>
#include <cstdio>

#include "vls_host.hpp"

namespace vlh {

double powiRef(double x, int m) {
  // the reference is C++98, where pow(double,int) is __builtin_powi: this
  // repeated-squaring product (libgcc __powidf2)
  unsigned n = m < 0 ? -(unsigned)m : (unsigned)m;
  double y = (n % 2) ? x : 1.0;
  while (n >>= 1) {
    x = x * x;
    if (n % 2) y = y * x;
  }
  return m < 0 ? 1.0 / y : y;
}

ScaleType scaleFromString(const std::string &s) {
  // scale::scaleTypeSelect (src/scale.hxx:22-32); names are lower case
  if (s == "linear") return LINEAR;
  if (s == "logarithmic") return LOGARITHMIC;
  if (s == "quadratic") return QUADRATIC;
  if (s == "cubic") return CUBIC;
  if (s == "quartic") return QUARTIC;
  throw Failure("ERROR: unknown scale type value: '" + s + "'");
}

std::string scaleToString(int s) {
  static const char *n[] = {"linear", "logarithmic", "quadratic", "cubic", "quartic"};
  if (s < 0 || s > 4) throw Failure("ERROR in typeselect::getString : no string associated to type 'scale type'");
  return n[s];
}

std::vector<double> genScale(double start, double stop, long N, int sct, bool cell) {
  std::vector<double> v;
  if (sct == LINEAR) {
    double dx = (stop - start) / (N);
    v.resize(N + 1);
    for (long i = 0; i < N + 1; i++) v[i] = start + dx * i;
  } else if (sct == LOGARITHMIC) {
    if (start * stop <= 0) throw Failure("error: cannot generate logscale from a 0-crossing range.");
    double s = (start < 0) ? -1 : 1;
    v = genScale(std::log(std::fabs(start)), std::log(std::fabs(stop)), N, LINEAR, false);
    for (long i = 0; i < N + 1; i++) v[i] = s * std::exp(v[i]);
  } else {
    int pp = sct;
    v = genScale(std::pow(start, 1. / pp), std::pow(stop, 1. / pp), N, LINEAR, false);
    for (long i = 0; i < N + 1; i++) v[i] = powiRef(v[i], pp);
  }
  if (cell) {
    for (size_t j = 0; j + 1 < v.size(); j++) v[j] = (v[j] + v[j + 1]) * 0.5;
    v.resize(v.size() - 1);
  } else {
    v.front() = start;
    v.back() = stop;
  }
  return v;
}

double rescale(double x, int sct) {
  if (sct == LOGARITHMIC) return std::log(x);
  if (sct == LINEAR) return x;
  return std::pow(x, 1. / sct);
}

void Grid21::init(const GridParams &gp) {
  p = gp;
  if (p.Uscale[0] != LINEAR)
    throw Failure("ERROR: the v_r axis must use a linear scale (grid.Uscale)");
  if (p.Pres < 3 || p.Ures[0] < 3 || p.Ures[1] < 1) throw Failure("ERROR: grid resolution too small");
  R_V = genScale(p.Pmin, p.Pmax, p.Pres, p.Pscale, false);
  R_Val = R_V;
  U_V = genScale(p.Umin[0], p.Umax[0], p.Ures[0], p.Uscale[0], false);
  U_Val = U_V;
  J_V = genScale(p.Umin[1], p.Umax[1], p.Ures[1], p.Uscale[1], false);
  J_Val = genScale(p.Umin[1], p.Umax[1], p.Ures[1], p.Uscale[1], true);
  nR = (int)R_Val.size();
  nU = (int)U_Val.size();
  nJ = (int)J_Val.size();
  nVal = (int64_t)nR * nU * nJ;
  dR.assign(nR, 0.0);
  dU.assign(nU, 0.0);
  dJ.assign(nJ, 0.0);
  for (int r = 0; r + 1 < nR; r++) dR[r] = R_V[r + 1] - R_V[r];
  for (int u = 0; u + 1 < nU; u++) dU[u] = U_V[u + 1] - U_V[u];
  // slice21Iterator::get_dJ tests j against the v vertex count
  // (src/slice21Iterator.hxx:86): slices with j >= nU-1 weigh nothing
  for (int j = 0; j < nJ; j++) dJ[j] = (j >= (int)U_V.size() - 1) ? 0 : (J_V[j + 1] - J_V[j]);
  // spline knots (vlasovSolver21_spline::setup :304-313; localSpline::init :60-93)
  int sc[2] = {p.Pscale, p.Uscale[0]};
  const std::vector<double> *V[2] = {&R_V, &U_V};
  long N[2] = {p.Pres, p.Ures[0]};
  for (int d = 0; d < 2; d++) {
    std::vector<double> scl = genScale(V[d]->front(), V[d]->back(), N[d], sc[d], false);
    double start = rescale(scl.front(), sc[d]);
    double stop = rescale(scl.back(), sc[d]);
    double step = (stop - start) / (double)(scl.size() - 1);
    splineStart[d] = start;
    splineHinv[d] = 1. / step;
  }
}

double Grid21::splineCoord(double pos, int dir) const {
  // localSpline::pos2Coord before its modf (src/localSpline.hxx:352-356)
  double q = rescale(pos, dir == 0 ? p.Pscale : p.Uscale[0]);
  return (q - splineStart[dir]) * splineHinv[dir];
}

vls_geometry Grid21::geometry() const {
  vls_geometry g{};
  g.nR = nR;
  g.nU = nU;
  g.nJ = nJ;
  g.R_Val = R_Val.data();
  g.U_Val = U_Val.data();
  g.J_Val = J_Val.data();
  g.dR = dR.data();
  g.dU = dU.data();
  g.dJ = dJ.data();
  for (int d = 0; d < 2; d++) {
    g.spline_start[d] = splineStart[d];
    g.spline_hinv[d] = splineHinv[d];
  }
  return g;
}

// ------------------------------------------------------------ generators

void InitGen::generate(const Grid21 &g, double *f, int nS) const {
  // species s of point i at f[i * nS + s] (gridBase::getDataPtr, src/grid_base.hxx:385)
#pragma omp parallel for schedule(static)
  for (int j = 0; j < g.nJ; j++)
    for (int u = 0; u < g.nU; u++)
      for (int r = 0; r < g.nR; r++)
        for (int s = 0; s < nS; s++)
          f[g.index(j, u, r) * nS + s] = valueAt(g.R_Val[r], g.U_Val[u], g.J_Val[j], s);
}

namespace {

// Uniform-density sphere with erf edge (initGen_UniformDensitySphere.hxx)
class UDF : public InitGen {
  // one value per species (initGen::MAX_SPECIES = 5, src/initGen.hxx:21)
  double M[5] = {}, R0[5] = {}, alpha[5] = {}, sr[5] = {}, gap0[5] = {}, gap1[5] = {};
  double rho0[5] = {}, C1[5] = {}, C2[5] = {};

 public:
  std::string type() const override { return "UDF"; }
  int maxSpecies() const override { return 5; }
  void setup(const Params &prm, const SolverParams &sp) override {
    // :43-74: value i of each parameter for species i (the gap's two values
    // at 2i, 2i+1)
    for (int i = 0; i < std::min(std::max(sp.nSpecies, 1), 5); i++) {
      R0[i] = prm.get<double>("R0", "init", 2.0, i);
      M[i] = prm.get<double>("M", "init", 1.0, i);
      alpha[i] = prm.get<double>("alpha", "init", 0.5, i);
      sr[i] = prm.get<double>("smoothingRadius", "init", 0.5, i);
      gap0[i] = prm.get<double>("gap", "init", 0.0, 2 * i);
      gap1[i] = prm.get<double>("gap", "init", 0.0, 2 * i + 1);
      rho0[i] = (3 * M[i]) / (4 * M_PI * powiRef(R0[i], 3));
      double sig0 = std::sqrt((alpha[i]) / (5 * R0[i]));
      C1[i] = rho0[i] * std::pow(2 * M_PI * sig0 * sig0, -1.5);
      C2[i] = 2 * sig0 * sig0;
    }
  }
  GridParams defaultGrid() const override {  // :82-102
    GridParams g;
    g.Pmin = 0.01; g.Pmax = 25;
    g.Umin[0] = -2.0; g.Umax[0] = 2.0; g.Umin[1] = 0; g.Umax[1] = 1.6;
    g.Pres = 200; g.Ures[0] = 200; g.Ures[1] = 200;
    g.Pscale = LOGARITHMIC; g.Uscale[0] = LINEAR; g.Uscale[1] = QUADRATIC;
    return g;
  }
  SolverParams defaultSolver() const override {  // :104-112
    SolverParams s;
    s.G = 1; s.T0 = 0; s.dt = 0.005; s.Tmax = 60.; s.nSpecies = 1;
    return s;
  }
  double valueAt(double r, double u, double j, int s) const override {  // computeF :158-173
    double sm = 0.5 * (1. + std::erf((R0[s] - r) / sr[s]));
    double res = sm * C1[s] * std::exp(-(u * u + (j * j) / (r * r)) / C2[s]);
    if (gap1[s] > 0) {
      double sm0 = 0.5 * (1. + std::erf((gap0[s] - r) / 0.15));
      double sm1 = 0.5 * (1. + std::erf(-(gap1[s] - r) / 0.15));
      res *= (sm0 + sm1);
    }
    if (res < 1.E-200) res = 1.E-200;
    return res;
  }
  // d/d(M, R0, alpha, smoothingRadius) of computeF: with z = (R0 - r)/s,
  // sm = (1 + erf z)/2 and E = C1 exp(-Q/C2) (Q = u^2 + j^2/r^2, times the
  // gap factor), f = sm E, dsm = exp(-z^2)/sqrt(pi) dz and
  // d ln E = d ln rho0 - 1.5 d ln sig0^2 + (Q/C2) d ln sig0^2, where
  // rho0 = 3M/(4 pi R0^3) and sig0^2 = alpha/(5 R0)
  bool dValueAt(double r, double u, double j, int s, const std::string &param, double &val,
                double &dval) const override {
    double dM = 0, dR0 = 0, dA = 0, dS = 0;
    if (param == "M") dM = 1;
    else if (param == "R0") dR0 = 1;
    else if (param == "alpha") dA = 1;
    else if (param == "smoothingRadius") dS = 1;
    else return false;
    val = valueAt(r, u, j, s);
    if (val <= 1.E-200) {
      dval = 0;
      return true;
    }
    const double z = (R0[s] - r) / sr[s];
    const double sm = 0.5 * (1. + std::erf(z));
    const double Q = u * u + (j * j) / (r * r);
    double E = C1[s] * std::exp(-Q / C2[s]);
    if (gap1[s] > 0) {
      double sm0 = 0.5 * (1. + std::erf((gap0[s] - r) / 0.15));
      double sm1 = 0.5 * (1. + std::erf(-(gap1[s] - r) / 0.15));
      E *= (sm0 + sm1);
    }
    const double dlnS = dA / alpha[s] - dR0 / R0[s];
    const double dlnE = dM / M[s] - 3 * dR0 / R0[s] - 1.5 * dlnS + Q / C2[s] * dlnS;
    const double dsm = std::exp(-z * z) / std::sqrt(M_PI) * (dR0 - z * dS) / sr[s];
    dval = dsm * E + sm * E * dlnE;
    return true;
  }
  double mass(int s) const override { return M[s]; }
};

// Power-law density sphere (initGen_PowerLawDensitySphere.hxx)
class PLDF : public InitGen {
  int nS = 1;
  double M[5] = {}, alpha[5] = {}, eta[5] = {}, n[5] = {}, sr[5] = {}, R0[5] = {};
  double rho0[5] = {}, sigO2[5] = {}, sigR2[5] = {}, C0[5] = {}, C1[5] = {}, C2[5] = {};
  bool R0given = false;  // init.R0 set; otherwise R0 = 1/eta follows eta

 public:
  std::string type() const override { return "PLDF"; }
  int maxSpecies() const override { return 5; }
  void setup(const Params &prm, const SolverParams &sp) override {
    // :44-80: value i of each parameter for species i
    nS = std::min(std::max(sp.nSpecies, 1), 5);
    for (int i = 0; i < nS; i++) {
      M[i] = prm.get<double>("M", "init", 1.0, i);
      alpha[i] = prm.get<double>("alpha", "init", 2.0, i);
      eta[i] = prm.get<double>("eta", "init", 0.1, i);
      n[i] = prm.get<double>("n", "init", 1.0, i);
      sr[i] = prm.get<double>("smoothingRadius", "init", 0.5, i);
      R0[i] = 1. / eta[i];
      R0[i] = prm.get<double>("R0", "init", R0[i], i);
      R0given = prm.isDefined("R0", "init");
      rho0[i] = ((3.0 - n[i]) * M[i]) / (4 * M_PI * std::pow(R0[i], 3.0 - n[i]));
      sigO2[i] = (3.0 - n[i]) / (5.0 - 2 * n[i]) * (M[i] * eta[i]) / (R0[i] * (alpha[i] * 0.5 + 1));
      sigR2[i] = sigO2[i] * 0.5 * alpha[i];
      C0[i] = rho0[i] / (std::sqrt(2 * sigR2[i] * M_PI * M_PI * M_PI) * sigO2[i]);
      C1[i] = 1.0 / (2 * sigR2[i]);
      C2[i] = 1.0 / (sigO2[i]);
    }
  }
  GridParams defaultGrid() const override {  // :88-121
    GridParams g;
    double sigma2 = 0;
    for (int i = 0; i < nS; i++) {
      double sig2 = sigO2[i] + 2 * sigR2[i];
      if (sigma2 < sig2) sigma2 = sig2;
    }
    g.Pmin = 0.01; g.Pmax = 200;
    g.Umin[0] = -16 * std::sqrt(sigma2); g.Umax[0] = 16 * std::sqrt(sigma2);
    g.Umin[1] = 0; g.Umax[1] = 1.6;
    g.Pres = 200; g.Ures[0] = 200; g.Ures[1] = 200;
    g.Pscale = LOGARITHMIC; g.Uscale[0] = LINEAR; g.Uscale[1] = QUADRATIC;
    return g;
  }
  SolverParams defaultSolver() const override {
    SolverParams s;
    s.G = 1; s.T0 = 0; s.dt = 0.005; s.Tmax = 60.; s.nSpecies = 1;
    return s;
  }
  double valueAt(double r, double u, double j, int s) const override {  // computeF :181-189
    double sm = 0.5 * (1. + std::erf((R0[s] - r) / sr[s]));
    double res =
        sm * C0[s] * std::exp(-u * u * C1[s] - ((j * j) / (r * r)) * C2[s]) * std::pow(r, -n[s]);
    if (res < 1.E-200) res = 1.E-200;
    return res;
  }
  // d/d(M, eta, alpha, R0, smoothingRadius) of computeF (eta moves R0 = 1/eta
  // too unless init.R0 is given): f = sm E, E = C0 exp(-u^2 C1 - (j/r)^2 C2)
  // r^-n, with d ln rho0 = dM/M - (3-n) dR0/R0,
  // d ln sigO2 = dM/M + deta/eta - dR0/R0 - (dalpha/2)/(alpha/2 + 1),
  // d ln sigR2 = d ln sigO2 + dalpha/alpha, d ln C0 = d ln rho0 - d ln sigR2/2
  // - d ln sigO2, dC1 = -C1 d ln sigR2, dC2 = -C2 d ln sigO2
  bool dValueAt(double r, double u, double j, int s, const std::string &param, double &val,
                double &dval) const override {
    double dM = 0, dEta = 0, dA = 0, dR0 = 0, dS = 0;
    if (param == "M") dM = 1;
    else if (param == "eta") {
      dEta = 1;
      if (!R0given) dR0 = -1. / (eta[s] * eta[s]);
    } else if (param == "alpha") dA = 1;
    else if (param == "R0") dR0 = 1;
    else if (param == "smoothingRadius") dS = 1;
    else return false;
    val = valueAt(r, u, j, s);
    if (val <= 1.E-200) {
      dval = 0;
      return true;
    }
    const double z = (R0[s] - r) / sr[s];
    const double sm = 0.5 * (1. + std::erf(z));
    const double jr2 = (j * j) / (r * r);
    const double E = C0[s] * std::exp(-u * u * C1[s] - jr2 * C2[s]) * std::pow(r, -n[s]);
    const double dlnRho0 = dM / M[s] - (3.0 - n[s]) * dR0 / R0[s];
    const double dlnSO = dM / M[s] + dEta / eta[s] - dR0 / R0[s] - (0.5 * dA) / (alpha[s] * 0.5 + 1);
    const double dlnSR = dlnSO + dA / alpha[s];
    const double dlnC0 = dlnRho0 - 0.5 * dlnSR - dlnSO;
    const double dlnE = dlnC0 + u * u * C1[s] * dlnSR + jr2 * C2[s] * dlnSO;
    const double dsm = std::exp(-z * z) / std::sqrt(M_PI) * (dR0 - z * dS) / sr[s];
    dval = dsm * E + sm * E * dlnE;
    return true;
  }
  double mass(int s) const override { return M[s]; }
};

// Hernquist sphere, isotropic or Osipkov-Merritt (initGen_Hernquist.hxx)
class Hernquist : public InitGen {
  double M = 1, R0 = 1, Ra = -1, G = 1, Vg = 1;

  double phi(double R) const { return -(G * M) / (R0 + R); }  // :92-94
  double F(double ph, double r, double u, double j) const {    // :133-154
    double eps = ph + (u * u + (j * j) / (r * r)) * 0.5;
    if (Ra > 0) eps = eps + (j * j) / (2 * Ra * Ra);
    if (eps >= 0) return 0;
    double q = std::sqrt(-(R0) / (G * M) * eps);
    double q2 = q * q;
    double A1 = (M) / (8. * std::sqrt(2.) * std::pow(M_PI * R0 * Vg, 3.)) * std::pow(1. - q2, -2.5);
    double A2 = 3 * std::asin(q) + q * std::sqrt(1. - q2) * (1. - 2 * q2) * (8 * q2 * q2 - 8 * q2 - 3);
    double result = A1 * A2;
    if (Ra > 0) {
      A1 = (M) / (std::sqrt(2.) * powiRef(M_PI * R0 * Vg, 3)) * powiRef(R0 / Ra, 2);
      A2 = q * (1. - 2 * q2);
      result += A1 * A2;
    }
    return result;
  }

 public:
  std::string type() const override { return "Hernquist"; }
  void setup(const Params &prm, const SolverParams &sp) override {  // :38-50
    G = sp.G;
    R0 = prm.get<double>("R0", "init", 1.0);
    M = prm.get<double>("M", "init", 1.0);
    Ra = prm.get<double>("Ra", "init", -1.0);
    Vg = std::sqrt((G * M) / (R0));
  }
  GridParams defaultGrid() const override {  // :58-75
    GridParams g;
    g.Pmin = 0.01; g.Pmax = 25 * 10;
    g.Umin[0] = -1.45; g.Umax[0] = 1.45; g.Umin[1] = 1.E-4; g.Umax[1] = 3.24 * 4;
    g.Pres = 45 * 2; g.Ures[0] = 40 * 2; g.Ures[1] = 20 * 4;
    g.Pscale = LOGARITHMIC; g.Uscale[0] = LINEAR; g.Uscale[1] = LOGARITHMIC;
    return g;
  }
  SolverParams defaultSolver() const override {  // :77-85
    SolverParams s;
    s.G = 1; s.T0 = 0; s.Tmax = 62; s.dt = 0.025; s.nSpecies = 1;
    return s;
  }
  double valueAt(double r, double u, double j, int) const override { return F(phi(r), r, u, j); }
  void generate(const Grid21 &g, double *f, int) const override {  // :163-204
    std::vector<double> Phi(g.nR);
    for (int r = 0; r < g.nR; r++) Phi[r] = phi(g.R_Val[r]);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < g.nJ; j++)
      for (int u = 0; u < g.nU; u++)
        for (int r = 0; r < g.nR; r++)
          f[g.index(j, u, r)] = F(Phi[r], g.R_Val[r], g.U_Val[u], g.J_Val[j]);
  }
  double mass(int) const override { return M; }
};

// Plummer sphere, isotropic (initGen_Plummer.hxx)
class Plummer : public InitGen {
  double M = 1, R0 = 1, G = 1;

  double phi(double R) const { return -(G * M) / std::sqrt(R * R + R0 * R0); }  // :91-93
  static double F(double ph, double r, double u, double j) {                    // :125-130
    double eps = ph + (u * u + (j * j) / (r * r)) * 0.5;
    if (eps >= 0) return 0;
    return std::pow(-2. * eps, 3.5) * (3. / (M_PI * M_PI * M_PI * 7.));
  }

 public:
  std::string type() const override { return "Plummer"; }
  void setup(const Params &prm, const SolverParams &sp) override {  // :37-44
    G = sp.G;
    R0 = prm.get<double>("R0", "init", 1.0);
    M = prm.get<double>("M", "init", 1.0);
  }
  GridParams defaultGrid() const override {  // :50-70
    GridParams g;
    g.Pmin = 0.01; g.Pmax = 25 * 4;
    g.Umin[0] = -1.45; g.Umax[0] = 1.45; g.Umin[1] = 0.00; g.Umax[1] = 3.24;
    g.Pres = 45 * 4; g.Ures[0] = 40 * 2; g.Ures[1] = 20 * 4;
    g.Pscale = LOGARITHMIC; g.Uscale[0] = LINEAR; g.Uscale[1] = QUADRATIC;
    return g;
  }
  SolverParams defaultSolver() const override {  // :72-80
    SolverParams s;
    s.G = 1; s.T0 = 0; s.Tmax = 62; s.dt = 0.1; s.nSpecies = 1;
    return s;
  }
  double valueAt(double r, double u, double j, int) const override { return F(phi(r), r, u, j); }
  void generate(const Grid21 &g, double *f, int) const override {  // :138-160
    std::vector<double> Phi(g.nR);
    for (int r = 0; r < g.nR; r++) Phi[r] = phi(g.R_Val[r]);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < g.nJ; j++)
      for (int u = 0; u < g.nU; u++)
        for (int r = 0; r < g.nR; r++)
          f[g.index(j, u, r)] = F(Phi[r], g.R_Val[r], g.U_Val[u], g.J_Val[j]);
  }
  double mass(int) const override { return M; }
};

}  // namespace

std::unique_ptr<InitGen> makeInitGen(const std::string &type) {
  // initGenAll::get (src/initGen_all.hxx:39-55)
  if (type == "UDF") return std::unique_ptr<InitGen>(new UDF());
  if (type == "PLDF") return std::unique_ptr<InitGen>(new PLDF());
  if (type == "Hernquist") return std::unique_ptr<InitGen>(new Hernquist());
  if (type == "Plummer") return std::unique_ptr<InitGen>(new Plummer());
  throw Failure("ERROR: unknown initial conditions type " + type + "\n.");
}

}  // namespace vlh
