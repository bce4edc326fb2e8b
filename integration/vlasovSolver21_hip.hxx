/*
 * vlasovSolver21_hip.hxx -- the reference-side binding of the MI355X step.
 *
 * A solver policy for vlasolve that keeps everything of
 * vlasovSolver21_spline (src/vlasovSolver21_spline.hxx) -- parameters,
 * initial conditions, renormalisation, static feet, delay lines,
 * statistics.txt, NDfield snapshots and .vls restarts -- and hands the time
 * step to libvlasolve_hip through the C ABI of include/vlasolve_hip.h.  It is
 * selected where the reference selects its solver (src/solver_traits.hxx:24-35,
 * src/setup.hxx:7), or used directly as in integration/main_hip.cpp, and is
 * compiled with the reference's own flags (C++98) against its unmodified
 * headers.
 *
 * The device state is built in beforeRunning() (src/vlasovSolver.hxx:211-214),
 * after the reference's init() has set up and renormalised f and the delay
 * lines (fresh runs and -reload alike).  makeOneStep() advances the device
 * only; f and the delay-line values come back to the host grid before every
 * snapshot, statistics line and restart, which the reference then writes as
 * it always does.  Errors print and exit(-1), as the reference does.
 */
#ifndef VLASOV_SOLVER21_HIP_HXX
#define VLASOV_SOLVER21_HIP_HXX

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "vlasovSolver21_spline.hxx"
#include "vlasolve_hip.h"

template <class gridHandlerType>
class vlasovSolver21_hip : public vlasovSolver21_spline<gridHandlerType> {
 public:
  typedef vlasovSolver21_spline<gridHandlerType> splT;
  typedef typename splT::b21T b21T;
  typedef typename splT::bT bT;
  typedef typename bT::mpiComT mpiComT;
  typedef typename bT::gridT gridT;
  typedef typename bT::dataT dataT;
  typedef typename bT::scaleT scaleT;
  typedef localSpline<2, dataT> coordSplineT;

  vlasovSolver21_hip(mpiComT &com) : splT(com), ctx(NULL), fStale(false), chainsStale(false) {}
  virtual ~vlasovSolver21_hip() { vls_destroy(ctx); }

  /* src/vlasovSolver21_spline.hxx:680-801, on the device */
  virtual void makeOneStep(double t) {
    (void)t;
    check(vls_step(ctx, 1), "vls_step");
    fStale = chainsStale = true;
  }

 protected:
  using splT::write;

  virtual void beforeRunning() {
    splT::beforeRunning();
    if (!ctx) upload();
  }

  virtual void statistics(bool reset) {
    pull(false);
    splT::statistics(reset);
  }

  virtual void snapshot() {
    pull(false);
    splT::snapshot();
  }

  virtual void write(FILE *f) {
    pull(true);
    splT::write(f);
  }

 private:
  vls_ctx *ctx;
  bool fStale, chainsStale;  // the device holds a newer f / newer delay-line values
  // the delay lines in the order they were uploaded: slice, point, cell count
  std::vector<long> chainSlice, chainPoint;
  std::vector<int> chainLen;

  static void check(int rc, const char *what) {
    if (rc) {
      const char *msg = vls_last_error();
      fprintf(stderr, "ERROR: %s\n", (msg && *msg) ? msg : what);
      exit(-1);
    }
  }

  // The spline's rescaled first knot and inverse knot spacing on one axis,
  // as localSpline::init computes them (src/localSpline.hxx:60-93) from the
  // vertices the reference's set-up passes it
  // (src/vlasovSolver21_spline.hxx:296-312).
  static void knots(const std::vector<double> &V, int sct, double &start, double &hinv) {
    double a = V.front(), b = V.back();
    scaleT::rescale(a, (typename scaleT::scaleTypeT)sct);
    scaleT::rescale(b, (typename scaleT::scaleTypeT)sct);
    const long dim = (long)V.size();
    start = a;
    hinv = 1. / ((b - a) / (dim - 1));
  }

  // The cells of one delay line, front first, through the chain's own
  // serialisation (cellChain::write, src/cellChain.hxx:90-96; dummyCell21,
  // src/dummyCell21_vect.hxx): a long count, then per cell R, U, an int
  // count and the values.
  struct Cell {
    double R, U;
    std::vector<double> val;
  };
  template <class chainT>
  static std::vector<Cell> cells_of(chainT &ch) {
    char *buf = NULL;
    size_t len = 0;
    FILE *f = open_memstream(&buf, &len);
    if (!f) { fprintf(stderr, "ERROR: open_memstream\n"); exit(-1); }
    ch.write(f);
    fclose(f);
    std::vector<Cell> out;
    const char *p = buf;
    long n = 0;
    memcpy(&n, p, sizeof(long));
    p += sizeof(long);
    for (long k = 0; k < n; k++) {
      Cell c;
      int nv = 0;
      memcpy(&c.R, p, sizeof(double)); p += sizeof(double);
      memcpy(&c.U, p, sizeof(double)); p += sizeof(double);
      memcpy(&nv, p, sizeof(int)); p += sizeof(int);
      c.val.resize(nv);
      if (nv) memcpy(&c.val[0], p, nv * sizeof(double));
      p += nv * sizeof(double);
      out.push_back(c);
    }
    free(buf);
    return out;
  }
  // ... and back (cellChain::read)
  template <class chainT>
  static void set_cells(chainT &ch, const std::vector<Cell> &cells) {
    char *buf = NULL;
    size_t len = 0;
    FILE *f = open_memstream(&buf, &len);
    if (!f) { fprintf(stderr, "ERROR: open_memstream\n"); exit(-1); }
    long n = (long)cells.size();
    fwrite(&n, sizeof(long), 1, f);
    for (size_t k = 0; k < cells.size(); k++) {
      int nv = (int)cells[k].val.size();
      fwrite(&cells[k].R, sizeof(double), 1, f);
      fwrite(&cells[k].U, sizeof(double), 1, f);
      fwrite(&nv, sizeof(int), 1, f);
      if (nv) fwrite(&cells[k].val[0], sizeof(double), nv, f);
    }
    fclose(f);
    FILE *g = fmemopen(buf, len, "r");
    if (!g) { fprintf(stderr, "ERROR: fmemopen\n"); exit(-1); }
    ch.read(g, false);
    fclose(g);
    free(buf);
  }

  void upload() {
    gridT *grid = bT::gh->getGrid();
    const int nR = (int)b21T::R_Val.size(), nU = (int)b21T::U_Val.size();
    const int nJ = (int)b21T::J_Val.size(), nS = bT::sp.nSpecies;
    const long nVal = (long)nR * nU * nJ;

    // quadrature widths (slice21Iterator::get_dR/dU/dJ, src/slice21Iterator.hxx:84-86,
    // get_dJ testing the v vertex count)
    std::vector<double> dR(nR, 0.0), dU(nU, 0.0), dJ(nJ, 0.0);
    for (int r = 0; r + 1 < nR; r++) dR[r] = b21T::R_V[r + 1] - b21T::R_V[r];
    for (int u = 0; u + 1 < nU; u++) dU[u] = b21T::U_V[u + 1] - b21T::U_V[u];
    for (int j = 0; j < nJ; j++)
      dJ[j] = (j >= (long)b21T::U_V.size() - 1) ? 0 : (b21T::J_V[j + 1] - b21T::J_V[j]);

    vls_geometry g;
    g.nR = nR;
    g.nU = nU;
    g.nJ = nJ;
    g.R_Val = &b21T::R_Val[0];
    g.U_Val = &b21T::U_Val[0];
    g.J_Val = &b21T::J_Val[0];
    g.dR = &dR[0];
    g.dU = &dU[0];
    g.dJ = &dJ[0];
    knots(b21T::R_V, (int)b21T::Rscale, g.spline_start[0], g.spline_hinv[0]);
    knots(b21T::U_V, (int)b21T::Uscale, g.spline_start[1], g.spline_hinv[1]);

    const bool delayed = b21T::kernelType == b21T::kernelTypeV::DELAYED;
    vls_solver_params sp;
    sp.G = bT::sp.G;
    sp.dt = bT::deltaTime;
    sp.no_linking = bT::noLinking ? 1 : 0;
    sp.kernel_type = delayed ? 0 : 1;
    sp.n_species = nS;
    check(vls_create(&g, &sp, 1, &ctx), "vls_create");
    // the grid's own layout: [j][u][r], species fastest (src/grid_base.hxx:385)
    check(vls_upload_f(ctx, 0, grid->getDataPtr()), "vls_upload_f");

    // static feet in spline units: the integer and fractional parts of
    // localSpline::pos2Coord (src/localSpline.hxx:352-358), as
    // decomposeDisplacement takes them (src/vlasovSolver21_spline.hxx:76-100)
    typename coordSplineT::initT ini;
    ini.start[0] = b21T::R_V.front();
    ini.stop[0] = b21T::R_V.back();
    ini.start[1] = b21T::U_V.front();
    ini.stop[1] = b21T::U_V.back();
    ini.N[0] = (int)b21T::R_C.size();
    ini.N[1] = (int)b21T::U_C.size();
    ini.scaleType[0] = b21T::Rscale;
    ini.scaleType[1] = b21T::Uscale;
    ini.vlType[0] = grid->valLocation_P;
    ini.vlType[1] = grid->valLocation_U;
    for (int a = 0; a < 2; a++)
      ini.bCond[a][0] = ini.bCond[a][1] = coordSplineT::boundaryTypeV::BT_NATURAL;
    coordSplineT spl(ini);
    std::vector<double> xr(nVal), xu(nVal);
    std::vector<uint8_t> flag(nVal);
    for (long i = 0; i < nVal; i++) {
      std::pair<int, double> a = spl.pos2Coord(b21T::disp_driftR[i], 0);
      std::pair<int, double> b = spl.pos2Coord(b21T::disp_driftU[i], 1);
      xr[i] = a.first + a.second;
      xu[i] = b.first + b.second;
      if (delayed && b21T::disp_driftR[i] < b21T::kernelRadius) flag[i] = VLS_FOOT_CHAIN;
      else if (b21T::disp_driftOut[i]) flag[i] = VLS_FOOT_OUT;
      else flag[i] = VLS_FOOT_EVAL;
    }
    check(vls_set_drift(ctx, 0, &xr[0], &xu[0], &flag[0]), "vls_set_drift");

    // delay lines (sphericalKernel, src/sphericalKernel.hxx), front first
    std::vector<int64_t> point;
    std::vector<int32_t> len, head;
    std::vector<double> bxr, bxu, vals;
    std::vector<uint8_t> bout;
    chainSlice.clear();
    chainPoint.clear();
    chainLen.clear();
    if (delayed)
      for (long j = 0; j < nJ; j++)
        for (typename b21T::kernelCellsItT it = b21T::kernel.begin(j); it != b21T::kernel.end(j);
             ++it) {
          std::vector<Cell> cells = cells_of(*it->second);
          const Cell &back = cells.back();
          point.push_back(it->first);
          len.push_back((int32_t)cells.size());
          head.push_back(0);
          std::pair<int, double> a = spl.pos2Coord(back.R, 0);
          std::pair<int, double> b = spl.pos2Coord(back.U, 1);
          bxr.push_back(a.first + a.second);
          bxu.push_back(b.first + b.second);
          bout.push_back((back.U <= b21T::U_Val.front()) || (back.U >= b21T::U_Val.back()));
          for (size_t k = 0; k < cells.size(); k++)
            for (int s = 0; s < nS; s++)
              vals.push_back(s < (int)cells[k].val.size() ? cells[k].val[s] : 0.0);
          chainSlice.push_back(j);
          chainPoint.push_back(it->first);
          chainLen.push_back((int)cells.size());
        }
    const int64_t n = (int64_t)point.size();
    check(vls_set_chains(ctx, 0, n, n ? &point[0] : NULL, n ? &len[0] : NULL,
                         n ? &head[0] : NULL, n ? &bxr[0] : NULL, n ? &bxu[0] : NULL,
                         n ? &bout[0] : NULL, n ? &vals[0] : NULL),
          "vls_set_chains");
    fStale = chainsStale = false;
  }

  // f (and, for restarts, the delay-line values) back into the host grid
  void pull(bool chains) {
    if (!ctx) return;
    if (fStale) {
      check(vls_download_f(ctx, 0, bT::gh->getGrid()->getDataPtr()), "vls_download_f");
      fStale = false;
    }
    if (chains && chainsStale && !chainPoint.empty()) {
      const int nS = bT::sp.nSpecies;
      size_t cellsTotal = 0;
      for (size_t c = 0; c < chainLen.size(); c++) cellsTotal += chainLen[c];
      std::vector<int32_t> head(chainLen.size());
      std::vector<double> vals(cellsTotal * nS);
      check(vls_get_chains(ctx, 0, &head[0], &vals[0]), "vls_get_chains");
      size_t off = 0;
      for (size_t c = 0; c < chainLen.size(); c++) {
        typename b21T::kernelCellsItT it = b21T::kernel.begin(chainSlice[c]);
        while (it != b21T::kernel.end(chainSlice[c]) && it->first != chainPoint[c]) ++it;
        std::vector<Cell> cells = cells_of(*it->second);
        const int L = chainLen[c];
        for (int k = 0; k < L; k++) {  // ring slot (head + k) mod L is cell k
          const size_t slot = off + (size_t)((head[c] + k) % L);
          for (int s = 0; s < nS && s < (int)cells[k].val.size(); s++)
            cells[k].val[s] = vals[slot * nS + s];
        }
        set_cells(*it->second, cells);
        off += L;
      }
    }
    if (chains) chainsStale = false;
  }
};

#endif
