// Host side of the drop-in solver surface: parameter parsing, grid geometry,
// initial conditions, static feet and delay lines, and the run loop that
// drives the device context.  Everything here runs once at set-up or between
// steps; every step itself runs on the GPU through the vls_* C ABI.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/vlasolve_hip.h"

namespace vlh {

// Raised for every user-facing error; the C ABI turns it into a status code
// and the vlasolve_hip executable prints it and exits like the reference.
struct Failure : std::runtime_error {
  int code;
  explicit Failure(const std::string &m, int c = VLS_ERR_ARG) : std::runtime_error(m), code(c) {}
};

extern int verbose;

// ------------------------------------------------------------ parameters
// Semantics of the reference's paramsParser (src/paramsParser.hxx):
// "-cat.key v1 v2 ..." on the command line, then an optional parameter file
// (first positional argument, default "params.ini") with "key v..." lines,
// '#' comments and the delimiters " =:{}[],\t".  Definitions from the command
// line come first, so they win.  Keys without a category go to "default".
class Params {
 public:
  Params() = default;
  Params(int argc, const char *const *argv);

  template <class T>
  T get(const std::string &key, const std::string &cat, const T &def, int which = -1) const;
  bool isDefined(const std::string &key, const std::string &cat) const;
  std::string reportUnused() const;
  const std::string &paramFile() const { return fileName_; }

 private:
  struct Value {
    std::vector<std::vector<std::string>> val;
    mutable std::vector<bool> wasRead;
  };
  std::map<std::string, std::map<std::string, Value>> cats_;
  std::string fileName_;

  void insert(const std::string &rawKey, const std::vector<std::string> &v);
  bool parseFile(const std::string &fname);
  const std::vector<std::string> *lookup(const std::string &key, const std::string &cat,
                                         int which, const Value **vout) const;
};

// ------------------------------------------------------------ geometry
enum ScaleType { LINEAR = 0, LOGARITHMIC = 1, QUADRATIC = 2, CUBIC = 3, QUARTIC = 4 };
ScaleType scaleFromString(const std::string &s);
std::string scaleToString(int s);

// node values of an axis (scale::genScale, src/scale.hxx:114-151)
std::vector<double> genScale(double start, double stop, long N, int sct, bool cell);
// coordinate used by the spline (scale::rescale, src/scale.hxx:175-189)
double rescale(double x, int sct);
// pow(double,int) of the reference's C++98 build: repeated squaring
double powiRef(double x, int n);

struct GridParams {
  double Pmin = 0, Pmax = 1;
  int Pres = 128, Pscale = LINEAR;
  double Umin[2] = {-1, -1}, Umax[2] = {1, 1};
  int Ures[2] = {128, 128}, Uscale[2] = {LINEAR, LINEAR};
};

// The (r, v_r, j) grid of one run: VERTEX r and v, CELL j
// (solver_traits.hxx:29-30; gridBase::init, src/grid_base.hxx:258-311).
struct Grid21 {
  GridParams p;
  int nR = 0, nU = 0, nJ = 0;
  int64_t nVal = 0;
  std::vector<double> R_Val, U_Val, J_Val, R_V, U_V, J_V;
  std::vector<double> dR, dU, dJ;  // quadrature widths (slice21Iterator.hxx:84-86)
  double splineStart[2] = {0, 0}, splineHinv[2] = {0, 0};

  void init(const GridParams &gp);
  int64_t index(int j, int u, int r) const { return ((int64_t)j * nU + u) * nR + r; }
  // spline-unit coordinate of a position on axis 0 (r) or 1 (v)
  double splineCoord(double p, int dir) const;
  vls_geometry geometry() const;
};

// ------------------------------------------------------------ initial conditions
struct SolverParams {
  double G = 1, T0 = 0, Tmax = 60, dt = 0.005;
  int nSpecies = 1;
  long snapshotEvery = 500, restartEvery = 5000, statsEvery = 25;
};

class InitGen {
 public:
  virtual ~InitGen() = default;
  virtual std::string type() const = 0;
  virtual void setup(const Params &prm, const SolverParams &sp) = 0;
  virtual GridParams defaultGrid() const = 0;
  virtual SolverParams defaultSolver() const = 0;
  // species s at (r, u, j) (valueAt(pos, s)); Hernquist and Plummer have one
  virtual double valueAt(double r, double u, double j, int s) const = 0;
  // fill f on the grid, species interleaved f[i * nS + s] (generate())
  virtual void generate(const Grid21 &g, double *f, int nS) const;
  virtual double mass(int s) const = 0;  // getMass(s)
  virtual int maxSpecies() const { return 1; }
  // valueAt and its derivative in the parameter `param` of species s ("M",
  // "R0", ...; the grid held fixed); false when the generator has no such
  // derivative.  val equals valueAt bit for bit; where valueAt returns its
  // floor the derivative is 0.
  virtual bool dValueAt(double r, double u, double j, int s, const std::string &param,
                        double &val, double &dval) const {
    (void)r; (void)u; (void)j; (void)s; (void)param; (void)val; (void)dval;
    return false;
  }
};

std::unique_ptr<InitGen> makeInitGen(const std::string &type);

// ------------------------------------------------------------ solver
enum KernelType { DELAYED = 0, REFLECTIVE = 1 };

// free streaming backward in time at fixed J (vlasovSolver21::advanceParticleDrift)
bool advanceDrift(double R, double U, double J, double dt, double R0, double &newR,
                  double &newU);

struct Chains {
  // chain c feeds point[c]; its cells, front first, are cellR/cellU[off[c]..off[c]+len[c]),
  // cell k's species values val[k * nSpecies .. (k + 1) * nSpecies)
  std::vector<int64_t> point, off;
  std::vector<int32_t> len, head;
  std::vector<double> cellR, cellU, val;
};

// Drop-in counterpart of vlasovSolver21_spline over the device context.
class Solver21 {
 public:
  Solver21() = default;
  ~Solver21();
  Solver21(const Solver21 &) = delete;
  Solver21 &operator=(const Solver21 &) = delete;

  // vlasovSolver::init (src/vlasovSolver.hxx:61-132); fname non-empty reloads a .vls.
  // With device=false only the host set-up runs (initial f, feet, delay lines).
  void init(const Params &prm, const std::string &reload, bool device = true);
  const std::vector<double> &hostF() const { return f_; }
  const std::vector<double> &driftR() const { return driftR_; }
  const std::vector<double> &driftU() const { return driftU_; }
  const std::vector<uint8_t> &driftOut() const { return driftOut_; }
  const Chains &chains() const { return chains_; }
  double kernelRadius() const { return kernelRadius_; }
  // Derivative of the initial state (f after the renormalisation, layout of
  // hostF(), and the delay-line values, layout of chains().val) in the
  // initial-condition parameter `param` of species s, the grid held fixed:
  // the derivative of valueAt through the renormalisation f <- f M / mass(f)
  // (vlasovSolver21::renormalize, src/vlasovSolver21.hxx:468-524).
  void initDerivative(const std::string &param, int s, std::vector<double> &df,
                      std::vector<double> &dvals) const;
  void run();                       // vlasovSolver::run
  void makeSteps(int n);            // n x makeOneStep, time and stepCount advance
  void statistics(bool reset);      // statistics.txt line
  void snapshot();                  // snapshot_%4.4f.ND
  void write(const std::string &fnameNoExt);  // .vls restart
  void stats(double out[7]);              // the species files' line
  void statsTotal(double out[7]);         // the statistics.txt line
  int nSpecies() const { return sp_.nSpecies; }
  void downloadF(double *f);
  void moments(double *dM, double *accel, double *mTot);

  const Grid21 &grid() const { return grid_; }
  double time() const { return time_; }
  long stepCount() const { return stepCount_; }
  vls_ctx *context() { return ctx_; }
  // device parameters of this run, and loading the host set-up (f, feet,
  // delay lines) into run `run` of a context whose geometry is grid().geometry()
  vls_solver_params deviceParams() const;
  void uploadTo(vls_ctx *ctx, int run) const;
  const SolverParams &solverParams() const { return sp_; }

 private:
  Grid21 grid_;
  SolverParams sp_;
  bool noLinking_ = false;
  int kernelType_ = DELAYED;
  int kernelBufferSize_ = 0;
  double kernelMass_ = 0, kernelRadius_ = 0;
  double time_ = 0;
  long stepCount_ = 0;
  std::vector<double> f_;         // host copy for set-up and I/O
  std::vector<double> driftR_, driftU_;
  std::vector<uint8_t> driftOut_;
  Chains chains_;
  vls_ctx *ctx_ = nullptr;
  std::string initType_;
  std::unique_ptr<InitGen> initGen_;  // the generator of the initial state (not on reloads)

  // quadrature21::integrate mass of species sp of an interleaved array of nS
  // species (the reference's single-thread order), as renormalize sums it
  double quadratureMass(const double *f, int nS, int sp) const;

  void setupFeet(std::unique_ptr<InitGen> *init);  // feet, chains (setup())
  void renormalize(double norm, int s);  // species s
  void upload();
  void readRestart(const std::string &fname, const Params &prm, bool device);
  bool checkTimer(long every) const { return every >= 0 && (stepCount_ % every) == 0; }
  bool checkTime() const;
};

// NDfield writer (Save_NDfield, src/NDfield.hxx:308-362)
void saveNDfield(const std::string &fname, const int *dims, int ndims, const double *x0,
                 const double *delta, const double *data, const std::string &comment);

}  // namespace vlh
