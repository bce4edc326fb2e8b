/*
 * vls_oracle.c -- CPU restatement of vlasolve's spherically symmetric
 * semi-Lagrangian spline time step (vlasovSolver21_spline::makeOneStep).
 *
 * TEST INFRASTRUCTURE ONLY: this file is the parity checker.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product
 * never links or calls it.
 *
 * Build with -O2 -ffp-contract=off and no -march flag: the reference is
 * compiled for baseline x86-64, which has no FMA, and every expression below
 * keeps the reference's evaluation order, so the results are bit-identical to
 * the reference's at the same reduction order.
 *
 * Pins (each function is checked by tests/test_oracle_*.py):
 *  - whole trajectories, feet, delay-line chains, scales and initial
 *    conditions: bit-identical to the reference executable built from its own
 *    sources (oracle/_ref, fixtures in tests/golden/) at nthreads=1;
 *  - spline build/evaluate: interpolation at nodes, exact reproduction of
 *    linear data, zero second derivative at natural ends, a dense
 *    brute-force solve of the same system, scipy's natural CubicSpline;
 *  - feet: conservation of H=(u^2+J^2/r^2)/2 and r*u - 2Ht, RK4 integration;
 *  - quadrature/statistics: closed forms for constant f, statistics.txt
 *    of the reference (10 significant digits).
 */
#include "vls_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>

/* slices per chunk of the device-order radial mass sum over j */
#define ORC_JCHUNK 20
#endif

static char g_err[512];
const char *orc_last_error(void) { return g_err; }
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }

/* ------------------------------------------------------------------ */
/* integer power with the reference's semantics.  The reference is C++98:
 * pow(double,int) resolves to __builtin_powi (libstdc++ <cmath>), i.e. the
 * repeated-squaring product below (libgcc __powidf2 / LLVM powi expansion).
 * Used where the reference writes pow(x, <int>), e.g. pow(R0,3)
 * (initGen_UniformDensitySphere.hxx:65) and pow(scale[i],pp) (scale.hxx:136). */
static double powi_ref(double x, int m)
{
  unsigned n = m < 0 ? -(unsigned)m : (unsigned)m;
  double y = (n % 2) ? x : 1.0;
  while (n >>= 1) {
    x = x * x;
    if (n % 2) y = y * x;
  }
  return m < 0 ? 1.0 / y : y;
}

/* scale::genScale, scale.hxx:114-151 */
long orc_gen_scale(double start, double stop, long N, int sct, int cell, double *out)
{
  long i;
  if (sct == ORC_LINEAR) {
    double dx = (stop - start) / (N);
    for (i = 0; i < N + 1; i++) out[i] = start + dx * i;
  } else if (sct == ORC_LOG) {
    double s = (start < 0) ? -1 : 1;
    orc_gen_scale(log(fabs(start)), log(fabs(stop)), N, ORC_LINEAR, 0, out);
    for (i = 0; i < N + 1; i++) out[i] = s * exp(out[i]);
  } else {
    int pp = sct;
    orc_gen_scale(pow(start, 1. / pp), pow(stop, 1. / pp), N, ORC_LINEAR, 0, out);
    for (i = 0; i < N + 1; i++) out[i] = powi_ref(out[i], pp);
  }
  if (cell) {
    for (i = 0; i < N; i++) out[i] = (out[i] + out[i + 1]) * 0.5;
    return N;
  }
  out[0] = start;
  out[N] = stop;
  return N + 1;
}

/* scale::rescale, scale.hxx:175-189 */
static double rescale(double x, int sct)
{
  if (sct == ORC_LOG) return log(x);
  if (sct == ORC_LINEAR) return x;
  return pow(x, 1. / sct);
}

/* ------------------------------------------------------------------ */
/* chain of delay cells (cellChain.hxx, dummyCell21_vect.hxx) */
typedef struct {
  long point; /* grid index i the chain feeds (map key, sphericalKernel.hxx:174-187) */
  long j;
  long n, cap;
  double *R, *U, *val; /* val[k * nS + s]: nS species values per cell */
  double *tval; /* tangent of val (tangent-linear model), or NULL */
} chain_t;

struct orc_state {
  orc_params p;
  long nR, nU, nJ, nVal;
  double *R_Val, *U_Val, *J_Val, *R_V, *U_V, *J_V;
  double *f;   /* nS planes of nVal points, species s at f + s * nVal */
  int nS;      /* species */
  /* initial-condition constants, per species */
  double C0[ORC_MAX_SPECIES], C1[ORC_MAX_SPECIES], C2[ORC_MAX_SPECIES], rho0[ORC_MAX_SPECIES];
  double R0, Vg;
  /* spline (localSpline.hxx:60-93) */
  int sdim[2];
  double sstart[2], sstep_inv[2];
  long stride1, cstride; /* padded row stride, padded slice size */
  double *coef;          /* nJ padded coefficient slices */
  double *tmpArr;        /* per-thread Thomas work rows */
  long tmpStride;
  double *colArr;        /* per-thread v-columns of the device-order kick */
  long colStride;
  /* feet (vlasovSolver21.hxx:92-94, vlasovSolver21_spline.hxx:69-72) */
  double *driftR, *driftU;
  unsigned char *driftOut;
  int *driftR_I, *driftU_I;
  double *driftR_F, *driftU_F;
  double kernelRadius;
  /* delay lines */
  long nChains;
  chain_t *chains;
  long *chainOfPoint; /* -1 or chain id */
  long *chainStart;   /* per slice j: first chain id (chains sorted by j) */
  /* last step's moments */
  double *dM, *Mr, mTot;
  /* tangent-linear model: tangents of f, of the coefficients and of the
   * acceleration (NULL until orc_tl_init) */
  double *tf, *tcoef, *tMr;
  double time;
  long stepCount;
};

/* ------------------------------------------------------------------ */
/* initial conditions */

/* UniformDensitySphere::setupParameters, initGen_UniformDensitySphere.hxx:65-74 */
static void udf_setup(orc_state *s, int k)
{
  const orc_params *p = &s->p;
  s->rho0[k] = (3 * p->spM[k]) / (4 * M_PI * powi_ref(p->spR0[k], 3));
  double sig0 = sqrt((p->spAlpha[k]) / (5 * p->spR0[k]));
  s->C1[k] = s->rho0[k] * pow(2 * M_PI * sig0 * sig0, -1.5);
  s->C2[k] = 2 * sig0 * sig0;
}

/* UniformDensitySphere::computeF, initGen_UniformDensitySphere.hxx:158-173 */
static double udf_F(const orc_state *s, int k, double r, double u, double j)
{
  const orc_params *p = &s->p;
  double sm = 0.5 * (1. + erf((p->spR0[k] - r) / p->spSR[k]));
  double res = sm * s->C1[k] * exp(-(u * u + (j * j) / (r * r)) / s->C2[k]);
  if (p->spGap1[k] > 0) {
    double sm0 = 0.5 * (1. + erf((p->spGap0[k] - r) / 0.15));
    double sm1 = 0.5 * (1. + erf(-(p->spGap1[k] - r) / 0.15));
    res *= (sm0 + sm1);
  }
  if (res < 1.E-200) res = 1.E-200;
  return res;
}

/* PowerLawDensitySphere::setupParameters, initGen_PowerLawDensitySphere.hxx:62-80
 * (R0 resolved by the caller: 1/eta unless given, :63-64) */
static void pldf_setup(orc_state *s, int k)
{
  const orc_params *p = &s->p;
  double n = p->spN[k], M = p->spM[k], R0 = p->spR0[k], alpha = p->spAlpha[k];
  s->rho0[k] = ((3.0 - n) * M) / (4 * M_PI * pow(R0, 3.0 - n));
  double sigO2 = (3.0 - n) / (5.0 - 2 * n) * (M * p->spEta[k]) / (R0 * (alpha * 0.5 + 1));
  double sigR2 = sigO2 * 0.5 * alpha;
  s->C0[k] = s->rho0[k] / (sqrt(2 * sigR2 * M_PI * M_PI * M_PI) * sigO2);
  s->C1[k] = 1.0 / (2 * sigR2);
  s->C2[k] = 1.0 / (sigO2);
}

/* PowerLawDensitySphere::computeF, initGen_PowerLawDensitySphere.hxx:181-189 */
static double pldf_F(const orc_state *s, int k, double r, double u, double j)
{
  const orc_params *p = &s->p;
  double sm = 0.5 * (1. + erf((p->spR0[k] - r) / p->spSR[k]));
  double res = sm * s->C0[k] * exp(-u * u * s->C1[k] - ((j * j) / (r * r)) * s->C2[k]) *
               pow(r, -p->spN[k]);
  if (res < 1.E-200) res = 1.E-200;
  return res;
}

/* Hernquist, initGen_Hernquist.hxx:38-50 (Vg), :92-94 (Phi), :133-154 (F) */
static double hq_phi(const orc_state *s, double R)
{
  return -(s->p.G * s->p.M) / (s->p.R0 + R);
}
static double hq_F(const orc_state *s, double phi, double r, double u, double j)
{
  const orc_params *p = &s->p;
  double eps = phi + (u * u + (j * j) / (r * r)) * 0.5;
  if (p->Ra > 0) eps = eps + (j * j) / (2 * p->Ra * p->Ra);
  if (eps >= 0) return 0;
  double q = sqrt(-(p->R0) / (p->G * p->M) * eps);
  double q2 = q * q;
  double A1 = (p->M) / (8. * sqrt(2.) * pow(M_PI * p->R0 * s->Vg, 3.)) * pow(1. - q2, -2.5);
  double A2 = 3 * asin(q) + q * sqrt(1. - q2) * (1. - 2 * q2) * (8 * q2 * q2 - 8 * q2 - 3);
  double result = A1 * A2;
  if (p->Ra > 0) {
    A1 = (p->M) / (sqrt(2.) * powi_ref(M_PI * p->R0 * s->Vg, 3)) * powi_ref(p->R0 / p->Ra, 2);
    A2 = q * (1. - 2 * q2);
    result += A1 * A2;
  }
  return result;
}

/* Plummer, initGen_Plummer.hxx:89-92 (Phi), :126-131 (F) */
static double pl_phi(const orc_state *s, double R)
{
  return -(s->p.G * s->p.M) / sqrt(R * R + s->p.R0 * s->p.R0);
}
static double pl_F(double phi, double r, double u, double j)
{
  double eps = phi + (u * u + (j * j) / (r * r)) * 0.5;
  if (eps >= 0) return 0;
  return pow(-2. * eps, 3.5) * (3. / (M_PI * M_PI * M_PI * 7.));
}

/* initGen::valueAt for each generator (UDF :175-179, PLDF :191-195, Hernquist :156-161,
 * Plummer initGen_Plummer.hxx:133-138) */
static double value_at(const orc_state *s, int k, double r, double u, double j)
{
  switch (s->p.initType) {
  case ORC_UDF: return udf_F(s, k, r, u, j);
  case ORC_PLDF: return pldf_F(s, k, r, u, j);
  case ORC_PLUMMER: return pl_F(pl_phi(s, r), r, u, j);
  default: return hq_F(s, hq_phi(s, r), r, u, j);
  }
}

/* ------------------------------------------------------------------ */
/* free streaming: vlasovSolver21::advanceParticleDrift, vlasovSolver21.hxx:328-408.
 * pow(x,2) in the reference is pow(double,int) = exact square x*x. */
static int advance(double R, double U, double J, double dt, double R0, double *oR, double *oU)
{
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
    tlim = sqrt(2 * R * R * H0 - J * J) / (2 * H0);
    if (R0 > 0) {
      double A, B, C, D;
      A = 2 * H0;
      B = -2 * sgn * sqrt(2 * R * R * H0 - J * J);
      C = (R * R - R0 * R0);
      D = B * B - 4 * A * C;
      if (D < 0) tR0 = 10 * dt;
      else tR0 = (-B - sqrt(D)) / (2 * A);
      if (tR0 < 0) {
        tR0 = (-B + sqrt(D)) / (2 * A);
        if (tR0 < 0) tR0 = 10 * dt;
      }
    } else tR0 = 10 * dt;
  }
  if ((tR0 < deltaT) && (tR0 < tlim)) {
    t = fabs(newR * newU) - 2 * sgn * H0 * tR0;
    newR = sqrt(t * t + J * J) * pow(2 * H0, -0.5);
    newU = (2 * H0 - (J * J) / (newR * newR));
    newU = -sqrt(newU);
    sgn = -1;
    H0 = 0.5 * (newU * newU + (J * J) / (newR * newR));
    deltaT -= tR0;
    tlim = 10 * dt;
    tR0 = 10 * dt;
  }
  if (tlim < deltaT) {
    t = fabs(newR * newU) - 2 * sgn * H0 * tlim;
    newR = sqrt(t * t + J * J) * pow(2 * H0, -0.5);
    newU = 0;
    H0 = 0.5 * (J * J) / (newR * newR);
    sgn = -1;
    deltaT -= tlim;
    tR0 -= tlim;
    tlim = 10 * dt;
    if (tR0 < deltaT) { /* the reference prints and exits here (:389-399) */
      set_err("advanceParticleDrift: pericentre and reflection in one step");
      return -1;
    }
  }
  t = fabs(newR * newU) - 2 * sgn * H0 * deltaT;
  newR = sqrt(t * t + J * J) * pow(2 * H0, -0.5);
  newU = (2 * H0 - (J * J) / (newR * newR));
  if (newU <= 0) newU = 0;
  else newU = sgn * sqrt(newU);
  *oR = newR;
  *oU = newU;
  return 0;
}

void orc_advance_drift(double R, double U, double J, double dt, double R0, double *newR,
                       double *newU)
{
  if (advance(R, U, J, dt, R0, newR, newU)) *newR = *newU = NAN;
}

/* ------------------------------------------------------------------ */
/* spline: localSpline<2>, natural boundary conditions on a single rank
 * (vlasovSolver21_spline.hxx:315-345) */

/* localSpline::create_1D, localSpline.hxx:953-1182, NATURAL rows (:976-981,
 * :1019-1024), elimination :1045-1059 and :1104-1112, last row :1143-1153,
 * back substitution :1155-1172.  A is a (M x 4) work array. */
static void create_1D(const orc_state *s, double *A, double *tab, int dir, long delta)
{
  const long M = (s->sdim[dir] + 2);
  const double hi = s->sstep_inv[dir];
  double *a = A;
  double *t = &tab[delta];
  long row;
  {
    double hi2 = hi * hi;
    a[0] = hi2; a[1] = -2.0 * hi2; a[2] = hi2; a[3] = 0;
  }
  a += 4;
  for (long i = 1; i < M - 1; i++, t += delta, a += 4) {
    a[0] = 1.0 / 6.0;
    a[1] = 4.0 / 6.0;
    a[2] = 1.0 / 6.0;
    a[3] = *t;
  }
  a = &A[(M - 1) * 4];
  {
    double hi2 = hi * hi;
    a[0] = hi2; a[1] = -2.0 * hi2; a[2] = hi2; a[3] = 0;
  }
  A[4 * (0) + 1] /= A[4 * (0) + 0];
  A[4 * (0) + 2] /= A[4 * (0) + 0];
  A[4 * (0) + 3] /= A[4 * (0) + 0];
  A[4 * (0) + 0] = 1.0;
  A[4 * (1) + 1] -= A[4 * (1) + 0] * A[4 * (0) + 1];
  A[4 * (1) + 2] -= A[4 * (1) + 0] * A[4 * (0) + 2];
  A[4 * (1) + 3] -= A[4 * (1) + 0] * A[4 * (0) + 3];
  A[4 * (1) + 0] = 0.0;
  A[4 * (1) + 2] /= A[4 * (1) + 1];
  A[4 * (1) + 3] /= A[4 * (1) + 1];
  A[4 * (1) + 1] = 1.0;
  row = 2;
  a = &A[4 * (row - 1) + 0];
  for (; row < M - 1; row++, a += 4) {
    a[4 + 1] -= a[4 + 0] * a[2];
    a[4 + 3] -= a[4 + 0] * a[3];
    a[4 + 2] /= a[4 + 1];
    a[4 + 3] /= a[4 + 1];
    a[4 + 0] = 0.0;
    a[4 + 1] = 1.0;
  }
  A[4 * (M - 1) + 1] -= A[4 * (M - 1) + 0] * A[4 * (M - 3) + 2];
  A[4 * (M - 1) + 3] -= A[4 * (M - 1) + 0] * A[4 * (M - 3) + 3];
  A[4 * (M - 1) + 2] -= A[4 * (M - 1) + 1] * A[4 * (M - 2) + 2];
  A[4 * (M - 1) + 3] -= A[4 * (M - 1) + 1] * A[4 * (M - 2) + 3];
  A[4 * (M - 1) + 3] /= A[4 * (M - 1) + 2];
  A[4 * (M - 1) + 2] = 1.0;

  t = &tab[delta * (M - 1)];
  *t = A[4 * (M - 1) + 3];
  t -= delta;
  for (row = M - 2; row > 0; row--, t -= delta) (*t) = A[4 * (row) + 3] - A[4 * (row) + 2] * t[delta];
  (*t) = A[4 * (0) + 3] - A[4 * (0) + 1] * t[delta] - A[4 * (0) + 2] * t[2 * delta];
}

/* localSpline::assign (:211-246) + create_2D (:1238-1258): the r-direction
 * solve of rows 1..dim1 (stride 1), then the v-direction solve of every padded
 * column (stride dim0+2).  data must hold (dim0+2)*(dim1+2) doubles. */
static void build_slice(const orc_state *s, double *A, const double *slice, double *data)
{
  const long d0 = s->sdim[0], d1 = s->sdim[1], st = s->stride1;
  long i, r, u;
  for (u = 0; u < d1; u++)
    for (r = 0; r < d0; r++) data[(u + 1) * st + (r + 1)] = slice[u * d0 + r];
  double *tab = &data[st];
  for (i = 0; i <= d1 - 1; i++, tab += st) create_1D(s, A, tab, 0, 1);
  tab = &data[0];
  for (i = -1; i <= d0; i++, tab += 1) create_1D(s, A, tab, 1, st);
}

static inline double BV2(double x) { return x * x * x; }
static inline double BV1(double x) { return 1.0 + 3.0 * x + 3.0 * x * x - 3.0 * x * x * x; }

/* localSpline::evaluateC(Nx,x,Ny,y), localSpline.hxx:381-398 */
static double evalC(long st, const double *data, int Nx, double x, int Ny, double y)
{
  Nx++;
  Ny++;
  const long ref[4] = {Nx + (Ny - 1) * st, Nx + Ny * st, Nx + (Ny + 1) * st, Nx + (Ny + 2) * st};
  const double b2xI = BV2(1. - x);
  const double b1xI = BV1(1. - x);
  const double b2x = BV2(x);
  const double b1x = BV1(x);
  double result = 1.0 / 36.0;
  result *=
      BV2(1. - y) * (data[ref[0] - 1] * b2xI + data[ref[0]] * b1xI + data[ref[0] + 1] * b1x + data[ref[0] + 2] * b2x) +
      BV1(1. - y) * (data[ref[1] - 1] * b2xI + data[ref[1]] * b1xI + data[ref[1] + 1] * b1x + data[ref[1] + 2] * b2x) +
      BV1(y) * (data[ref[2] - 1] * b2xI + data[ref[2]] * b1xI + data[ref[2] + 1] * b1x + data[ref[2] + 2] * b2x) +
      BV2(y) * (data[ref[3] - 1] * b2xI + data[ref[3]] * b1xI + data[ref[3] + 1] * b1x + data[ref[3] + 2] * b2x);
  return result;
}

/* localSpline::evaluateNxPy, localSpline.hxx:423-441 */
static double evalNxPy(const orc_state *s, const double *data, int Nx, double yy)
{
  const long st = s->stride1;
  Nx++;
  yy = rescale(yy, s->p.Uscale[0]);
  double y = (yy - s->sstart[1]) * s->sstep_inv[1];
  y = modf(y, &yy);
  int Ny = 1 + (int)yy;
  const long ref[4] = {Nx + (Ny - 1) * st, Nx + Ny * st, Nx + (Ny + 1) * st, Nx + (Ny + 2) * st};
  double result = 1.0 / 36.0;
  result *= BV2(1. - y) * (data[ref[0] - 1] + data[ref[0]] * 4 + data[ref[0] + 1]) +
            BV1(1. - y) * (data[ref[1] - 1] + data[ref[1]] * 4 + data[ref[1] + 1]) +
            BV1(y) * (data[ref[2] - 1] + data[ref[2]] * 4 + data[ref[2] + 1]) +
            BV2(y) * (data[ref[3] - 1] + data[ref[3]] * 4 + data[ref[3] + 1]);
  return result;
}

/* localSpline::pos2Coord, localSpline.hxx:352-358 */
static void pos2coord(const orc_state *s, double p, int dir, int *I, double *F)
{
  p = rescale(p, dir == 0 ? s->p.Pscale : s->p.Uscale[0]);
  double x = (p - s->sstart[dir]) * s->sstep_inv[dir];
  x = modf(x, &p);
  *I = (int)p;
  *F = x;
}

/* localSpline::evaluate(x,y), localSpline.hxx:477-488 */
static double eval_xy(const orc_state *s, const double *data, double xx, double yy)
{
  int I, J;
  double x, y;
  pos2coord(s, xx, 0, &I, &x);
  pos2coord(s, yy, 1, &J, &y);
  return evalC(s->stride1, data, I, x, J, y);
}

/* ------------------------------------------------------------------ */
/* quadrature (quadrature21.hxx, integrand21.hxx, slice21Iterator.hxx:84-94) */

static inline double get_dR(const orc_state *s, long r)
{
  return (r >= s->nR - 1) ? 0 : (s->R_V[r + 1] - s->R_V[r]);
}
static inline double get_dU(const orc_state *s, long u)
{
  return (u >= s->nU - 1) ? 0 : (s->U_V[u + 1] - s->U_V[u]);
}
/* slice21Iterator.hxx:86 compares j against U_V.size(), not J_V.size() */
static inline double get_dJ(const orc_state *s, long j)
{
  return (j >= s->nU - 1) ? 0 : (s->J_V[j + 1] - s->J_V[j]);
}

/* slice21Iterator::get_avg, slice21Iterator.hxx:90-94 */
static inline double avg4(const orc_state *s, const double *f, long i)
{
  return ((f[i]) + f[i + 1] + f[i + s->nR] + f[i + s->nR + 1]) * 0.25;
}

/* the integrand of int21::f(0, ns) (integrand21.hxx:21-39): get_avg() of the
 * one field, or with several fields 0 + get_avg(0) + get_avg(1) + ...; the
 * planes of f are nVal apart */
static inline double avgS(const orc_state *s, const double *f, long i, int ns)
{
  if (ns <= 1) return avg4(s, f, i);
  double v = 0;
  for (int k = 0; k < ns; k++) v += avg4(s, f + (size_t)k * s->nVal, i);
  return v;
}

/* static schedule chunk of iteration range [0,n) for thread t of T
 * (contiguous, balanced; identical to every OpenMP runtime when T | n) */
static void chunk(long n, int T, int t, long *lo, long *hi)
{
  long q = n / T, rem = n % T;
  *lo = t * q + (t < rem ? t : rem);
  *hi = *lo + q + (t < rem ? 1 : 0);
}

/* per-point moment term of integrate1D (quadrature21.hxx:111-119) */
static inline double moment_term(const orc_state *s, const double *f, long j, long u, long r,
                                 double Ct1, int ns)
{
  double dU = get_dU(s, u);
  if (dU <= 0) return NAN;
  double dR = get_dR(s, r);
  if (dR <= 0) return NAN;
  return Ct1 * dU * dR * avgS(s, f, (j * s->nU + u) * s->nR + r, ns);
}

/* the device's form of one species' term: ((Ct1 dU 0.25) dR) (sum of the
 * four corners in get_avg's order), get_avg's factor 0.25 folded into the
 * first product.  Scaling by 0.25 is exact, so this is the value of
 * moment_term whenever the corners' sum times 0.25 is (sums above 2^-1020):
 * both round the same product.  Several species keep the reference's form. */
static inline double moment_term_dev(const orc_state *s, const double *f, long j, long u, long r,
                                     double Ct1, int ns)
{
  if (ns > 1) return moment_term(s, f, j, u, r, Ct1, ns);
  double dU = get_dU(s, u);
  if (dU <= 0) return NAN;
  double dR = get_dR(s, r);
  if (dR <= 0) return NAN;
  const long i = (j * s->nU + u) * s->nR + r;
  const double sum = ((f[i] + f[i + 1]) + f[i + s->nR]) + f[i + s->nR + 1];
  return ((Ct1 * dU * 0.25) * dR) * sum;
}

/* quadrature21::integrate1D with int21::f(0, ns) (quadrature21.hxx:88-142) */
static void integrate1D(orc_state *s, const double *f, double *Mrv, int ns)
{
  const long nR = s->nR, nU = s->nU, nJ = s->nJ;
  long r;
  int T = s->p.nthreads;
  if (T >= 1) {
    double *tmp = calloc((size_t)T * nR, sizeof(double));
#pragma omp parallel for schedule(static, 1)
    for (int t = 0; t < T; t++) {
      long lo, hi;
      chunk(nJ, T, t, &lo, &hi);
      double *Mr = tmp + (size_t)t * nR;
      for (long j = lo; j < hi; j++) {
        double Ct1 = 8 * M_PI * M_PI * s->J_Val[j] * get_dJ(s, j);
        for (long u = 0; u < nU; u++)
          for (long rr = 0; rr < nR; rr++) {
            double dU = get_dU(s, u);
            if (dU <= 0) continue;
            double dR = get_dR(s, rr);
            if (dR <= 0) continue;
            Mr[rr] += Ct1 * dU * dR * avgS(s, f, (j * nU + u) * nR + rr, ns);
          }
      }
    }
    for (r = 0; r < nR; r++) {
      Mrv[r] = 0;
      for (int t = 0; t < T; t++) Mrv[r] += tmp[(size_t)t * nR + r];
    }
    free(tmp);
    return;
  }
  /* device order: the rows u of a slice form waves of 64 and groups of 16
   * (u = 64w + 16q + t); G_wq = serial sum over the group's rows,
   * W_w = ((G_w0 + G_w1) + G_w2) + G_w3, P[j][r] = ((W_0 + W_1) + ...) over
   * the waves; the slices form chunks of ORC_JCHUNK (j = 20c + i),
   * S_c = serial sum over the chunk's slices, dM[r] = serial sum of the S_c
   * (like the reference's per-thread partials, quadrature21.hxx:104-133) */
  double *P = calloc((size_t)nJ * nR, sizeof(double));
  const long nW = (nU + 63) / 64;
#pragma omp parallel for schedule(static)
  for (long j = 0; j < nJ; j++) {
    double Ct1 = 8 * M_PI * M_PI * s->J_Val[j] * get_dJ(s, j);
    for (long rr = 0; rr < nR - 1; rr++) {
      double Pj = 0.0;
      for (long w = 0; w < nW; w++) {
        double G[4];
        for (int q = 0; q < 4; q++) {
          double acc = 0.0;
          for (long t = 0; t < 16; t++) {
            long u = 64 * w + 16 * q + t;
            if (u >= nU - 1) continue;
            double tval = moment_term_dev(s, f, j, u, rr, Ct1, ns);
            if (tval == tval) acc += tval;
          }
          G[q] = acc;
        }
        double W = ((G[0] + G[1]) + G[2]) + G[3];
        Pj = w == 0 ? W : Pj + W;
      }
      P[(size_t)j * nR + rr] = Pj;
    }
  }
  for (r = 0; r < nR; r++) {
    double acc = 0;
    for (long c0 = 0; c0 < nJ; c0 += ORC_JCHUNK) {
      double S = 0.0;
      for (long j = c0; j < nJ && j < c0 + ORC_JCHUNK; j++) S += P[(size_t)j * nR + r];
      acc += S;
    }
    Mrv[r] = acc;
  }
  free(P);
}

/* integrand kinds of integrand21.hxx: f (:66-79), kineticEnergy (:82-95),
 * entropy (:98-112), L1Norm (:115-128), L2Norm (:131-145) */
enum { IG_F = 0, IG_EK, IG_S, IG_L1, IG_L2 };

static inline double integrand(int kind, double r, double u, double j, double v)
{
  switch (kind) {
  case IG_F: return v;
  case IG_EK: return 0.5 * v * (u * u + j * j / (r * r));
  case IG_S: return (v <= 0) ? 0 : -v * log(v);
  case IG_L1: return fabs(v);
  default: return v * v;
  }
}

/* quadrature21::integrate (quadrature21.hxx:57-85), integrand over the sum
 * of the first ns fields */
static double integrate(orc_state *s, const double *f, int kind, int ns)
{
  const long nR = s->nR, nU = s->nU, nJ = s->nJ;
  int T = s->p.nthreads;
  if (T >= 1) {
    double *res = calloc((size_t)T, sizeof(double));
#pragma omp parallel for schedule(static, 1)
    for (int t = 0; t < T; t++) {
      long lo, hi;
      chunk(nJ, T, t, &lo, &hi);
      double acc = 0;
      for (long j = lo; j < hi; j++) {
        double J = s->J_Val[j];
        double dJ = get_dJ(s, j);
        double Ct1 = 8 * M_PI * M_PI * J * dJ;
        for (long u = 0; u < nU; u++)
          for (long r = 0; r < nR; r++) {
            double dU = get_dU(s, u);
            if (dU <= 0) continue;
            double dR = get_dR(s, r);
            if (dR <= 0) continue;
            double v = avgS(s, f, (j * nU + u) * nR + r, ns);
            acc += Ct1 * dU * dR * integrand(kind, s->R_Val[r] + 0.5 * dR, s->U_Val[u] + 0.5 * dU, J, v);
          }
      }
      res[t] = acc;
    }
    for (int t = 1; t < T; t++) res[0] += res[t];
    double out = res[0];
    free(res);
    return out;
  }
  /* device order: per (j,r) serial over u, per j serial over r, then over j */
  double *Q = calloc((size_t)nJ, sizeof(double));
#pragma omp parallel for schedule(static)
  for (long j = 0; j < nJ; j++) {
    double J = s->J_Val[j];
    double dJ = get_dJ(s, j);
    double Ct1 = 8 * M_PI * M_PI * J * dJ;
    double qj = 0;
    for (long r = 0; r < nR - 1; r++) {
      double dR = get_dR(s, r);
      double acc = 0;
      for (long u = 0; u < nU - 1; u++) {
        double dU = get_dU(s, u);
        double v = avgS(s, f, (j * nU + u) * nR + r, ns);
        acc += Ct1 * dU * dR * integrand(kind, s->R_Val[r] + 0.5 * dR, s->U_Val[u] + 0.5 * dU, J, v);
      }
      qj += acc;
    }
    Q[j] = qj;
  }
  double tot = 0;
  for (long j = 0; j < nJ; j++) tot += Q[j];
  free(Q);
  return tot;
}

/* gridHandler::accumulate with addSpare=1 (grid_handler.hxx:350-407):
 * exclusive prefix Mr[i] = sum_{k<i} dM[k]; returns the total */
static double accumulate(const double *dM, double *Mr, long n)
{
  double prev = 0; /* result[0] = 0 */
  for (long i = 0; i < n; i++) {
    Mr[i] = prev;
    prev += dM[i]; /* result[i+1] = result[i] + dM[i] */
  }
  return prev;
}

/* ------------------------------------------------------------------ */
/* append a cell with its nS species values (v NULL: zeros) */
static void chain_push(chain_t *c, double R, double U, const double *v, int nS)
{
  if (c->n == c->cap) {
    c->cap = c->cap ? 2 * c->cap : 4;
    c->R = realloc(c->R, c->cap * sizeof(double));
    c->U = realloc(c->U, c->cap * sizeof(double));
    c->val = realloc(c->val, c->cap * nS * sizeof(double));
  }
  c->R[c->n] = R;
  c->U[c->n] = U;
  for (int k = 0; k < nS; k++) c->val[c->n * nS + k] = v ? v[k] : 0;
  c->n++;
}

/* sphericalKernel::readAt (sphericalKernel.hxx:48-54) with cellChain::roll
 * (cellChain.hxx:49-56): read species k of the front cell; with roll,
 * successive swaps of the cells' value vectors move every cell's values one
 * place towards the front and the old front's to the back */
static double chain_read(chain_t *c, int k, int nS, int roll)
{
  double result = c->val[k];
  if (roll)
    for (long q = 0; q + 1 < c->n; q++)
      for (int m = 0; m < nS; m++) {
        double t = c->val[q * nS + m];
        c->val[q * nS + m] = c->val[(q + 1) * nS + m];
        c->val[(q + 1) * nS + m] = t;
      }
  return result;
}

/* vlasovSolver21::generateDisplacement (vlasovSolver21.hxx:102-143) */
static int gen_displacement(orc_state *s, double dt, int reflective)
{
  const double dth = 0.5 * dt;
  const double R0 = reflective ? s->kernelRadius : -1;
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (long j = 0; j < s->nJ; j++) {
    double J = s->J_Val[j];
    for (long u = 0; u < s->nU; u++)
      for (long r = 0; r < s->nR; r++) {
        long i = (j * s->nU + u) * s->nR + r;
        double nr, nu;
        if (advance(s->R_Val[r], s->U_Val[u], J, dth, R0, &nr, &nu)) bad = 1;
        s->driftR[i] = nr;
        s->driftU[i] = nu;
        s->driftOut[i] = (nr >= s->R_Val[s->nR - 1]) || (nr <= s->R_Val[0]) ||
                         (nu >= s->U_Val[s->nU - 1]) || (nu <= s->U_Val[0]);
      }
  }
  return bad ? -1 : 0;
}

/* vlasovSolver21::setupBoundaryConditions (vlasovSolver21.hxx:265-326) */
static int setup_chains(orc_state *s, double dt)
{
  const double dth = 0.5 * dt;
  long cap = 0;
  s->nChains = 0;
  s->chains = NULL;
  for (long j = 0; j < s->nJ; j++) {
    double J = s->J_Val[j];
    s->chainStart[j] = s->nChains;
    for (long u = 0; u < s->nU; u++)
      for (long r = 0; r < s->nR; r++) {
        long i = (j * s->nU + u) * s->nR + r;
        if (!(s->driftR[i] < s->kernelRadius)) continue;
        if (s->nChains == cap) {
          cap = cap ? 2 * cap : 64;
          s->chains = realloc(s->chains, cap * sizeof(chain_t));
        }
        chain_t *c = &s->chains[s->nChains];
        memset(c, 0, sizeof *c);
        c->point = i;
        c->j = j;
        s->chainOfPoint[i] = s->nChains++;
        double deltaT = dth, nr, nu;
        if (advance(s->R_Val[r], s->U_Val[u], J, deltaT, -1, &nr, &nu)) return -1;
        while (nr < s->kernelRadius) {
          double v[ORC_MAX_SPECIES];
          for (int k = 0; k < s->nS; k++) v[k] = value_at(s, k, nr, nu, J);
          chain_push(c, nr, nu, v, s->nS);
          deltaT += dth;
          if (advance(s->R_Val[r], s->U_Val[u], J, deltaT, -1, &nr, &nu)) return -1;
        }
        /* the back cell starts at 0 for every species (the reference makes it
         * with one value, dummyCell21(R,U,1), vlasovSolver21.hxx:322: with
         * several species its later writes and reads of species >= 1 fall
         * outside that cell's vector) */
        chain_push(c, nr, nu, NULL, s->nS);
      }
  }
  s->chainStart[s->nJ] = s->nChains;
  return 0;
}

/* vlasovSolver21::renormalize(norm, which) (vlasovSolver21.hxx:468-524) for
 * species k, kernelMass = 0 */
static void renormalize(orc_state *s, double norm, int k)
{
  if (norm <= 0) return;
  /* set-up runs on the host in the reference's order; in the device order
   * mode the serial (one-thread) order is used */
  int T = s->p.nthreads;
  if (T < 1) s->p.nthreads = 1;
  double *fk = s->f + (size_t)k * s->nVal;
  double mass = integrate(s, fk, IG_F, 1) + 0.0;
  s->p.nthreads = T;
  double fac = norm / mass;
  for (long i = 0; i < s->nVal; i++) fk[i] *= fac;
  for (long c = 0; c < s->nChains; c++)
    for (long q = 0; q < s->chains[c].n; q++) s->chains[c].val[q * s->nS + k] *= fac;
}

/* ------------------------------------------------------------------ */
static int thread_count(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
static int thread_id(void)
{
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

/* the splines of every slice of the plane f (setupSplines for one species) */
static void build_all(orc_state *s, const double *f)
{
#pragma omp parallel for schedule(static)
  for (long j = 0; j < s->nJ; j++)
    build_slice(s, s->tmpArr + (size_t)thread_id() * s->tmpStride, f + j * s->nU * s->nR,
                s->coef + (size_t)j * s->cstride);
}

/* kick loop of one species (plane f), vlasovSolver21_spline.hxx:750-765 */
static void kick(orc_state *s, double *f)
{
  const double Ulo = s->U_Val[0], Uhi = s->U_Val[s->nU - 1];
#pragma omp parallel for schedule(static)
  for (long j = 0; j < s->nJ; j++) {
    const double *c = s->coef + (size_t)j * s->cstride;
    for (long u = 0; u < s->nU; u++)
      for (long r = 0; r < s->nR; r++) {
        long i = (j * s->nU + u) * s->nR + r;
        double newU = s->U_Val[u] + s->Mr[r];
        if ((newU <= Ulo) || (newU >= Uhi)) f[i] = 0;
        else f[i] = evalNxPy(s, c, (int)r, newU);
      }
  }
}

/* Device-order kick (nthreads == 0).  The kick evaluates the 2D spline on r
 * nodes only, where its r-direction combination (c[r-1] + 4c[r] + c[r+1])/6
 * gives back the r-solve's right-hand sides exactly: the (1,4,1)/6
 * combination of the 2D coefficients is the natural v-spline of the column
 * of f, and evaluateNxPy's (1/36) sum_k B_k(t) (c + 4c + c)[n+k] equals
 * (1/6) sum_k B_k(t) d[n+k] with d the v-spline coefficients of the column.
 * The device computes that form (one v-solve per column, no r-solve); it
 * agrees with the reference's to rounding.  With tf (tangent-linear model)
 * the tangent column goes through the same solve and the target's motion
 * adds (hinv/6) sum_k B_k'(t) d[n+k] * tMr[r]. */
static void kick_cols(orc_state *s, double *f, double *tf)
{
  const long nR = s->nR, nU = s->nU, M = s->sdim[1] + 2;
  const double Ulo = s->U_Val[0], Uhi = s->U_Val[nU - 1];
  const double st = s->sstart[1], hinv = s->sstep_inv[1];
#pragma omp parallel for schedule(static)
  for (long j = 0; j < s->nJ; j++) {
    double *A = s->tmpArr + (size_t)thread_id() * s->tmpStride;
    double *d = s->colArr + (size_t)thread_id() * s->colStride, *td = d + (M + 1);
    for (long r = 0; r < nR; r++) {
      for (long u = 0; u < nU; u++) {
        d[u + 1] = f[(j * nU + u) * nR + r];
        if (tf) td[u + 1] = tf[(j * nU + u) * nR + r];
      }
      create_1D(s, A, d, 1, 1);
      d[M] = 0; /* the row past the column, read with weight BV2(0) = 0 */
      if (tf) {
        create_1D(s, A, td, 1, 1);
        td[M] = 0;
      }
      for (long u = 0; u < nU; u++) {
        const long i = (j * nU + u) * nR + r;
        const double newU = s->U_Val[u] + s->Mr[r];
        if ((newU <= Ulo) || (newU >= Uhi)) {
          f[i] = 0;
          if (tf) tf[i] = 0;
          continue;
        }
        double yy = rescale(newU, s->p.Uscale[0]);
        double y = (yy - st) * hinv;
        y = modf(y, &yy);
        const long n = (long)yy;
        double result = 1.0 / 6.0;
        result *= BV2(1. - y) * d[n] + BV1(1. - y) * d[n + 1] + BV1(y) * d[n + 2] + BV2(y) * d[n + 3];
        f[i] = result;
        if (tf) {
          const double w = 1 - y;
          const double b0 = -3 * w * w, b1 = -(3 + 6 * w - 9 * w * w), b2 = 3 + 6 * y - 9 * y * y,
                       b3 = 3 * y * y;
          double tr = 1.0 / 6.0;
          tr *= BV2(1. - y) * td[n] + BV1(1. - y) * td[n + 1] + BV1(y) * td[n + 2] + BV2(y) * td[n + 3];
          tf[i] = tr + hinv / 6.0 * (b0 * d[n] + b1 * d[n + 1] + b2 * d[n + 2] + b3 * d[n + 3]) * s->tMr[r];
        }
      }
    }
  }
}

/* drift loop of species k, vlasovSolver21_spline.hxx:769-798 (delayed-kernel
 * back values :776-786, then per point :788-796); the delay lines turn on the
 * last species' reads (:793) */
static void drift(orc_state *s, int k)
{
  const double Ulo = s->U_Val[0], Uhi = s->U_Val[s->nU - 1];
  const int nS = s->nS, roll = k == nS - 1;
  double *f = s->f + (size_t)k * s->nVal;
#pragma omp parallel for schedule(static)
  for (long j = 0; j < s->nJ; j++) {
    const double *c = s->coef + (size_t)j * s->cstride;
    if (s->p.kernelType == ORC_DELAYED)
      for (long q = s->chainStart[j]; q < s->chainStart[j + 1]; q++) {
        chain_t *ch = &s->chains[q];
        double R = ch->R[ch->n - 1], U = ch->U[ch->n - 1];
        double *bv = &ch->val[(ch->n - 1) * nS + k];
        if ((U <= Ulo) || (U >= Uhi)) *bv = 0;
        else *bv = eval_xy(s, c, R, U);
      }
    for (long u = 0; u < s->nU; u++)
      for (long r = 0; r < s->nR; r++) {
        long i = (j * s->nU + u) * s->nR + r;
        if (s->driftR[i] < s->kernelRadius) f[i] = chain_read(&s->chains[s->chainOfPoint[i]], k, nS, roll);
        else if (s->driftOut[i]) f[i] = 0;
        else f[i] = evalC(s->stride1, c, s->driftR_I[i], s->driftR_F[i], s->driftU_I[i], s->driftU_F[i]);
      }
  }
}

static void moment_and_accel(orc_state *s)
{
  integrate1D(s, s->f, s->dM, s->nS);
  s->mTot = accumulate(s->dM, s->Mr, s->nR);
  for (long i = 0; i < s->nR; i++) s->Mr[i] *= s->p.G / (s->R_Val[i] * s->R_Val[i]) * s->p.dt;
}

/* vlasovSolver21_spline::makeOneStep, vlasovSolver21_spline.hxx:680-801: every
 * phase species by species (each species' splines are independent) */
void orc_step(orc_state *s, int nsteps)
{
  for (int n = 0; n < nsteps; n++) {
    if (s->p.noLinking)
      for (int k = 0; k < s->nS; k++) {
        build_all(s, s->f + (size_t)k * s->nVal);
        drift(s, k);
      }
    moment_and_accel(s);
    for (int k = 0; k < s->nS; k++) {
      if (s->p.nthreads == 0) {
        kick_cols(s, s->f + (size_t)k * s->nVal, NULL);
      } else {
        build_all(s, s->f + (size_t)k * s->nVal);
        kick(s, s->f + (size_t)k * s->nVal);
      }
    }
    for (int k = 0; k < s->nS; k++) {
      build_all(s, s->f + (size_t)k * s->nVal);
      drift(s, k);
    }
    s->time += s->p.dt;
    s->stepCount++;
  }
}

/* ------------------------------------------------------------------ */
orc_state *orc_create(const orc_params *p)
{
  orc_state *s = calloc(1, sizeof *s);
  s->p = *p;
  if (p->Uscale[0] != ORC_LINEAR) {
    set_err("only a linear v_r scale is supported");
    free(s);
    return NULL;
  }
  s->nR = p->Pres + 1;
  s->nU = p->Ures[0] + 1;
  s->nJ = p->Ures[1];
  s->nVal = s->nR * s->nU * s->nJ;
  s->R_Val = malloc((s->nR + 1) * sizeof(double));
  s->R_V = malloc((s->nR + 1) * sizeof(double));
  s->U_Val = malloc((s->nU + 1) * sizeof(double));
  s->U_V = malloc((s->nU + 1) * sizeof(double));
  s->J_Val = malloc((s->nJ + 2) * sizeof(double));
  s->J_V = malloc((s->nJ + 2) * sizeof(double));
  /* gridBase::init, grid_base.hxx:271-311 (VERTEX r and v, CELL j) */
  orc_gen_scale(p->Pmin, p->Pmax, p->Pres, p->Pscale, 0, s->R_Val);
  orc_gen_scale(p->Pmin, p->Pmax, p->Pres, p->Pscale, 0, s->R_V);
  orc_gen_scale(p->Umin[0], p->Umax[0], p->Ures[0], p->Uscale[0], 0, s->U_Val);
  orc_gen_scale(p->Umin[0], p->Umax[0], p->Ures[0], p->Uscale[0], 0, s->U_V);
  orc_gen_scale(p->Umin[1], p->Umax[1], p->Ures[1], p->Uscale[1], 1, s->J_Val);
  orc_gen_scale(p->Umin[1], p->Umax[1], p->Ures[1], p->Uscale[1], 0, s->J_V);

  s->nS = p->nSpecies < 1 ? 1 : p->nSpecies;
  if (s->nS > ORC_MAX_SPECIES ||
      (s->nS > 1 && p->initType != ORC_UDF && p->initType != ORC_PLDF)) {
    set_err("several species need UDF or PLDF initial conditions (at most 5 species)");
    free(s);
    return NULL;
  }
  s->f = calloc((size_t)s->nVal * s->nS, sizeof(double));
  for (int k = 0; k < s->nS; k++) {
    if (p->initType == ORC_UDF) udf_setup(s, k);
    else if (p->initType == ORC_PLDF) pldf_setup(s, k);
  }
  if (p->initType != ORC_UDF && p->initType != ORC_PLDF) s->Vg = sqrt((p->G * p->M) / (p->R0));
  /* generate (UDF :182-208, PLDF :197-224, Hernquist :163-204, Plummer
   * initGen_Plummer.hxx:141-162: Phi at the r nodes, as valueAt) */
  for (int k = 0; k < s->nS; k++) {
    double *fk = s->f + (size_t)k * s->nVal;
#pragma omp parallel for schedule(static)
    for (long j = 0; j < s->nJ; j++) {
      double J = s->J_Val[j];
      for (long u = 0; u < s->nU; u++)
        for (long r = 0; r < s->nR; r++) {
          long i = (j * s->nU + u) * s->nR + r;
          if (p->initType == ORC_HERNQUIST) fk[i] = hq_F(s, hq_phi(s, s->R_Val[r]), s->R_Val[r], s->U_Val[u], J);
          else fk[i] = value_at(s, k, s->R_Val[r], s->U_Val[u], J);
        }
    }
  }

  /* spline set-up, vlasovSolver21_spline.hxx:296-314 + localSpline::init :60-93 */
  {
    double tmp[2][8];
    int sc[2] = {p->Pscale, p->Uscale[0]};
    double st[2] = {s->R_V[0], s->U_V[0]}, sp_[2] = {s->R_V[s->nR - 1], s->U_V[s->nU - 1]};
    long N[2] = {p->Pres, p->Ures[0]};
    (void)tmp;
    for (int d = 0; d < 2; d++) {
      double *scl = malloc((N[d] + 2) * sizeof(double));
      long n = orc_gen_scale(st[d], sp_[d], N[d], sc[d], 0, scl);
      s->sdim[d] = (int)n;
      s->sstart[d] = rescale(scl[0], sc[d]);
      double stop = rescale(scl[n - 1], sc[d]);
      double step = (stop - s->sstart[d]) / (n - 1);
      s->sstep_inv[d] = 1. / step;
      free(scl);
    }
  }
  s->stride1 = s->sdim[0] + 2;
  s->cstride = (long)(s->sdim[0] + 2) * (s->sdim[1] + 2);
  /* a foot on the last knot makes the stencil read one column past its row
   * (and one row past its slice) with weight zero; pad the tail so those
   * reads stay inside the allocation */
  s->coef = calloc((size_t)s->nJ * s->cstride + 2 * (size_t)(s->sdim[0] + 2) + 8, sizeof(double));
  s->tmpStride = 4 * (2 + (s->sdim[0] > s->sdim[1] ? s->sdim[0] : s->sdim[1]));
  s->tmpArr = malloc((size_t)thread_count() * s->tmpStride * sizeof(double));
  s->colStride = 2 * ((long)s->sdim[1] + 3);
  s->colArr = malloc((size_t)thread_count() * s->colStride * sizeof(double));

  /* vlasovSolver21::setup, vlasovSolver21.hxx:411-446 */
  s->kernelRadius = s->R_Val[p->kernelBufferSize];
  s->driftR = malloc(s->nVal * sizeof(double));
  s->driftU = malloc(s->nVal * sizeof(double));
  s->driftOut = malloc(s->nVal);
  s->chainOfPoint = malloc(s->nVal * sizeof(long));
  s->chainStart = calloc(s->nJ + 1, sizeof(long));
  for (long i = 0; i < s->nVal; i++) s->chainOfPoint[i] = -1;
  double dtFeet = p->noLinking ? p->dt : p->dt * 2;
  if (gen_displacement(s, dtFeet, p->kernelType == ORC_REFLECTIVE)) goto fail;
  /* vlasovSolver21::setup(init,params), :533-548 */
  if (p->kernelType == ORC_DELAYED) {
    if (setup_chains(s, dtFeet)) goto fail;
  } else {
    for (long i = 0; i < s->nVal; i++)
      if (s->driftR[i] < s->kernelRadius) {
        set_err("reflective kernel: a foot falls inside the kernel radius");
        goto fail;
      }
  }
  /* decomposeDisplacement, vlasovSolver21_spline.hxx:76-100 */
  s->driftR_I = malloc(s->nVal * sizeof(int));
  s->driftU_I = malloc(s->nVal * sizeof(int));
  s->driftR_F = malloc(s->nVal * sizeof(double));
  s->driftU_F = malloc(s->nVal * sizeof(double));
  for (long i = 0; i < s->nVal; i++) {
    pos2coord(s, s->driftR[i], 0, &s->driftR_I[i], &s->driftR_F[i]);
    pos2coord(s, s->driftU[i], 1, &s->driftU_I[i], &s->driftU_F[i]);
  }
  /* vlasovSolver::init renormalizes each species to its mass (:94-95) */
  for (int k = 0; k < s->nS; k++)
    renormalize(s, p->initType == ORC_UDF || p->initType == ORC_PLDF ? p->spM[k] : p->M, k);
  s->dM = calloc(s->nR, sizeof(double));
  s->Mr = calloc(s->nR, sizeof(double));
  return s;
fail:
  orc_destroy(s);
  return NULL;
}

/* ------------------------------------------------------------------ */
/* Tangent-linear model of the step (linking, delayed or reflective kernel):
 * the derivative of the discrete map in a direction (df, dvals).  Every
 * phase of the step is linear in (f, chain values) except the kick, whose
 * target moves with the acceleration a(f); its tangent adds
 * d(evaluateNxPy)/dnewU * da.  Not part of the reference: this is the check
 * for the adjoint of the HIP path. */

/* d/dnewU of evalNxPy for the linear v scale: hinv/36 * sum B_k'(t) G_k */
static double evalNxPy_dU(const orc_state *s, const double *data, int Nx, double yy)
{
  const long st = s->stride1;
  Nx++;
  double y = (yy - s->sstart[1]) * s->sstep_inv[1];
  y = modf(y, &yy);
  int Ny = 1 + (int)yy;
  const long ref[4] = {Nx + (Ny - 1) * st, Nx + Ny * st, Nx + (Ny + 1) * st, Nx + (Ny + 2) * st};
  double G[4];
  for (int k = 0; k < 4; k++) G[k] = data[ref[k] - 1] + data[ref[k]] * 4 + data[ref[k] + 1];
  const double w = 1 - y;
  const double d0 = -3 * w * w, d1 = -(3 + 6 * w - 9 * w * w), d2 = 3 + 6 * y - 9 * y * y,
               d3 = 3 * y * y;
  return s->sstep_inv[1] / 36.0 * (d0 * G[0] + d1 * G[1] + d2 * G[2] + d3 * G[3]);
}

/* overwrite the delay-line values, cells front first in chain order */
/* cell values in the layout of orc_get_chains (nS per cell) */
void orc_set_chain_vals(orc_state *s, const double *vals)
{
  long off = 0;
  for (long c = 0; c < s->nChains; c++) {
    memcpy(s->chains[c].val, vals + off * s->nS, s->chains[c].n * s->nS * sizeof(double));
    off += s->chains[c].n;
  }
}

int orc_tl_init(orc_state *s, const double *df, const double *dvals)
{
  if (s->p.noLinking || s->p.Uscale[0] != ORC_LINEAR || s->nS != 1) {
    set_err("tangent-linear model: linking steps, a linear v scale and one species only");
    return -1;
  }
  if (!s->tf) {
    s->tf = malloc(s->nVal * sizeof(double));
    s->tcoef = calloc((size_t)s->nJ * s->cstride + 2 * (size_t)(s->sdim[0] + 2) + 8, sizeof(double));
    s->tMr = malloc(s->nR * sizeof(double));
  }
  memcpy(s->tf, df, s->nVal * sizeof(double));
  long off = 0;
  for (long c = 0; c < s->nChains; c++) {
    chain_t *ch = &s->chains[c];
    free(ch->tval);
    ch->tval = malloc(ch->n * sizeof(double));
    for (long k = 0; k < ch->n; k++) ch->tval[k] = dvals ? dvals[off + k] : 0.0;
    off += ch->n;
  }
  return 0;
}

void orc_tl_get(const orc_state *s, double *df) { memcpy(df, s->tf, s->nVal * sizeof(double)); }

/* exclusive-scan mass profile M(<R_r) of f and its tangent */
void orc_tl_mass(orc_state *s, double *M, double *dM)
{
  double *tmp = malloc(s->nR * sizeof(double));
  integrate1D(s, s->f, tmp, 1);
  accumulate(tmp, M, s->nR);
  integrate1D(s, s->tf, tmp, 1);
  accumulate(tmp, dM, s->nR);
  free(tmp);
}

static void tl_build(orc_state *s)
{
#pragma omp parallel for schedule(static)
  for (long j = 0; j < s->nJ; j++)
    build_slice(s, s->tmpArr + (size_t)thread_id() * s->tmpStride, s->tf + j * s->nU * s->nR,
                s->tcoef + (size_t)j * s->cstride);
}

static void tl_step1(orc_state *s)
{
  const long nR = s->nR, nU = s->nU, nJ = s->nJ;
  /* moment and acceleration, and their tangents */
  moment_and_accel(s);
  double *tdM = malloc(nR * sizeof(double));
  integrate1D(s, s->tf, tdM, 1);
  accumulate(tdM, s->tMr, nR);
  for (long i = 0; i < nR; i++) s->tMr[i] *= s->p.G / (s->R_Val[i] * s->R_Val[i]) * s->p.dt;
  free(tdM);
  /* kick */
  const double Ulo = s->U_Val[0], Uhi = s->U_Val[nU - 1];
  if (s->p.nthreads == 0) {
    kick_cols(s, s->f, s->tf);
  } else {
    build_all(s, s->f);
    tl_build(s);
  }
#pragma omp parallel for schedule(static)
  for (long j = 0; j < (s->p.nthreads == 0 ? 0 : nJ); j++) {
    const double *c = s->coef + (size_t)j * s->cstride;
    const double *tc = s->tcoef + (size_t)j * s->cstride;
    for (long u = 0; u < nU; u++)
      for (long r = 0; r < nR; r++) {
        long i = (j * nU + u) * nR + r;
        double newU = s->U_Val[u] + s->Mr[r];
        if ((newU <= Ulo) || (newU >= Uhi)) {
          s->f[i] = 0;
          s->tf[i] = 0;
        } else {
          s->f[i] = evalNxPy(s, c, (int)r, newU);
          s->tf[i] = evalNxPy(s, tc, (int)r, newU) + evalNxPy_dU(s, c, (int)r, newU) * s->tMr[r];
        }
      }
  }
  /* drift with the delay lines */
  build_all(s, s->f);
  tl_build(s);
#pragma omp parallel for schedule(static)
  for (long j = 0; j < nJ; j++) {
    const double *c = s->coef + (size_t)j * s->cstride;
    const double *tc = s->tcoef + (size_t)j * s->cstride;
    if (s->p.kernelType == ORC_DELAYED)
      for (long k = s->chainStart[j]; k < s->chainStart[j + 1]; k++) {
        chain_t *ch = &s->chains[k];
        double R = ch->R[ch->n - 1], U = ch->U[ch->n - 1];
        int out = (U <= Ulo) || (U >= Uhi);
        ch->val[ch->n - 1] = out ? 0 : eval_xy(s, c, R, U);
        ch->tval[ch->n - 1] = out ? 0 : eval_xy(s, tc, R, U);
      }
    for (long u = 0; u < nU; u++)
      for (long r = 0; r < nR; r++) {
        long i = (j * nU + u) * nR + r;
        if (s->driftR[i] < s->kernelRadius) {
          chain_t *ch = &s->chains[s->chainOfPoint[i]];
          double t0 = ch->tval[0];
          for (long k = 0; k + 1 < ch->n; k++) ch->tval[k] = ch->tval[k + 1];
          ch->tval[ch->n - 1] = t0;
          s->tf[i] = t0;
          s->f[i] = chain_read(ch, 0, 1, 1);
        } else if (s->driftOut[i]) {
          s->f[i] = 0;
          s->tf[i] = 0;
        } else {
          s->f[i] = evalC(s->stride1, c, s->driftR_I[i], s->driftR_F[i], s->driftU_I[i], s->driftU_F[i]);
          s->tf[i] = evalC(s->stride1, tc, s->driftR_I[i], s->driftR_F[i], s->driftU_I[i], s->driftU_F[i]);
        }
      }
  }
  s->time += s->p.dt;
  s->stepCount++;
}

int orc_tl_step(orc_state *s, int nsteps)
{
  if (!s->tf) {
    set_err("orc_tl_step: call orc_tl_init first");
    return -1;
  }
  for (int n = 0; n < nsteps; n++) tl_step1(s);
  return 0;
}

void orc_destroy(orc_state *s)
{
  if (!s) return;
  for (long c = 0; c < s->nChains; c++) {
    free(s->chains[c].R);
    free(s->chains[c].U);
    free(s->chains[c].val);
    free(s->chains[c].tval);
  }
  free(s->chains);
  free(s->R_Val); free(s->R_V); free(s->U_Val); free(s->U_V); free(s->J_Val); free(s->J_V);
  free(s->f); free(s->coef); free(s->tmpArr); free(s->colArr);
  free(s->driftR); free(s->driftU); free(s->driftOut);
  free(s->driftR_I); free(s->driftU_I); free(s->driftR_F); free(s->driftU_F);
  free(s->chainOfPoint); free(s->chainStart); free(s->dM); free(s->Mr);
  free(s->tf); free(s->tcoef); free(s->tMr);
  free(s);
}

void orc_dims(const orc_state *s, long dims[3])
{
  dims[0] = s->nR;
  dims[1] = s->nU;
  dims[2] = s->nJ;
}
long orc_nval(const orc_state *s) { return s->nVal; }
int orc_nspecies(const orc_state *s) { return s->nS; }
/* the planes <-> the reference's interleaved fields (gridBase::getDataPtr,
 * grid_base.hxx:385: arr[i * nFields + s]) */
void orc_get_f(const orc_state *s, double *out)
{
  for (int k = 0; k < s->nS; k++)
    for (long i = 0; i < s->nVal; i++) out[i * s->nS + k] = s->f[(size_t)k * s->nVal + i];
}
void orc_set_f(orc_state *s, const double *in)
{
  for (int k = 0; k < s->nS; k++)
    for (long i = 0; i < s->nVal; i++) s->f[(size_t)k * s->nVal + i] = in[i * s->nS + k];
}
double orc_kernel_radius(const orc_state *s) { return s->kernelRadius; }

void orc_get_axes(const orc_state *s, double *R_Val, double *U_Val, double *J_Val, double *J_V)
{
  if (R_Val) memcpy(R_Val, s->R_Val, s->nR * sizeof(double));
  if (U_Val) memcpy(U_Val, s->U_Val, s->nU * sizeof(double));
  if (J_Val) memcpy(J_Val, s->J_Val, s->nJ * sizeof(double));
  if (J_V) memcpy(J_V, s->J_V, (s->nJ + 1) * sizeof(double));
}

void orc_last_moments(const orc_state *s, double *dM, double *accel, double *mTot)
{
  if (dM) memcpy(dM, s->dM, s->nR * sizeof(double));
  if (accel) memcpy(accel, s->Mr, s->nR * sizeof(double));
  if (mTot) *mTot = s->mTot;
}

void orc_moments(orc_state *s, double *dM, double *Mexcl, double *mTot)
{
  double *d = malloc(s->nR * sizeof(double));
  double *m = malloc(s->nR * sizeof(double));
  integrate1D(s, s->f, d, s->nS);
  double t = accumulate(d, m, s->nR);
  if (dM) memcpy(dM, d, s->nR * sizeof(double));
  if (Mexcl) memcpy(Mexcl, m, s->nR * sizeof(double));
  if (mTot) *mTot = t;
  free(d);
  free(m);
}

/* vlasovSolver21::statistics + potentialEnergy (vlasovSolver21.hxx:627-689).
 * potentialEnergy gathers integrate1D with addSpare=1 (quadrature21.hxx:36,
 * grid_handler.hxx:350-372): its array is {0, dM[0], ..., dM[nR-1]}, so shell
 * k (between R[k] and R[k+1]) meets the mass below it, M(<R[k]), at its own
 * midpoint.  The last iteration (index nR) reads one past R but multiplies
 * dM[nR-1] = 0 (the last node has no width), so it adds nothing. */
void orc_stats(orc_state *s, double out[7])
{
  double *dM = malloc(s->nR * sizeof(double));
  integrate1D(s, s->f, dM, s->nS); /* int21::f(0, nSpecies) */
  double M = 0.0 + 0.0, Ep = 0; /* Mr[0] + kernelMass */
  for (long i = 1; i < s->nR; i++) {
    double r = 0.5 * (s->R_Val[i - 1] + s->R_Val[i]);
    Ep -= (dM[i - 1] * s->p.G * M) / r;
    M += dM[i - 1];
  }
  M += dM[s->nR - 1];
  free(dM);
  /* the default integrands read the first field only (get_avg(),
   * slice21Iterator.hxx:90-94), whatever the number of species */
  double Ek = integrate(s, s->f, IG_EK, 1);
  double E = Ek + Ep;
  double S = integrate(s, s->f, IG_S, 1);
  double L1 = integrate(s, s->f, IG_L1, 1);
  double L2 = integrate(s, s->f, IG_L2, 1);
  out[0] = M; out[1] = E; out[2] = Ek; out[3] = Ep; out[4] = S; out[5] = L1; out[6] = L2;
}

/* the statistics.txt line with several species (vlasovSolver21.hxx:656-710):
 * the species loop adds the species line nSpecies times, and S and L2 are
 * those of the summed fields (int21::entropy(0,nsp), L2Norm(0,nsp)) */
void orc_stats_total(orc_state *s, double out[7])
{
  double line[7];
  orc_stats(s, line);
  if (s->nS == 1) {
    memcpy(out, line, sizeof line);
    return;
  }
  double Ep_tot = 0, M_tot = 0, Ek_tot = 0, E_tot = 0, L1_tot = 0;
  for (int k = 0; k < s->nS; k++) {
    Ep_tot += line[3];
    M_tot += line[0];
    Ek_tot += line[2];
    E_tot += line[1];
    L1_tot += line[5];
  }
  out[0] = M_tot; out[1] = E_tot; out[2] = Ek_tot; out[3] = Ep_tot;
  out[4] = integrate(s, s->f, IG_S, s->nS);
  out[5] = L1_tot;
  out[6] = integrate(s, s->f, IG_L2, s->nS);
}

void orc_get_feet(const orc_state *s, double *driftR, double *driftU, uint8_t *flag)
{
  for (long i = 0; i < s->nVal; i++) {
    if (driftR) driftR[i] = s->driftR[i];
    if (driftU) driftU[i] = s->driftU[i];
    if (flag) flag[i] = (s->driftR[i] < s->kernelRadius) ? 2 : (s->driftOut[i] ? 1 : 0);
  }
}

long orc_num_chains(const orc_state *s) { return s->nChains; }
long orc_chain_cells(const orc_state *s)
{
  long n = 0;
  for (long c = 0; c < s->nChains; c++) n += s->chains[c].n;
  return n;
}

void orc_get_chains(const orc_state *s, int64_t *point, int64_t *off, int32_t *len, double *cellR,
                    double *cellU, double *cellVal)
{
  long o = 0;
  for (long c = 0; c < s->nChains; c++) {
    const chain_t *ch = &s->chains[c];
    if (point) point[c] = ch->point;
    if (off) off[c] = o;
    if (len) len[c] = (int32_t)ch->n;
    for (long k = 0; k < ch->n; k++) {
      if (cellR) cellR[o + k] = ch->R[k];
      if (cellU) cellU[o + k] = ch->U[k];
      if (cellVal)
        for (int m = 0; m < s->nS; m++) cellVal[(o + k) * s->nS + m] = ch->val[k * s->nS + m];
    }
    o += ch->n;
  }
}

void orc_spline_build(const orc_state *s, const double *slice, double *coef)
{
  double *A = malloc(s->tmpStride * sizeof(double));
  memset(coef, 0, s->cstride * sizeof(double)); /* caller provides cstride doubles */
  build_slice(s, A, slice, coef);
  free(A);
}
double orc_spline_evalC(const orc_state *s, const double *coef, int Nx, double x, int Ny, double y)
{
  return evalC(s->stride1, coef, Nx, x, Ny, y);
}
double orc_spline_evalNxPy(const orc_state *s, const double *coef, int Nx, double yy)
{
  return evalNxPy(s, coef, Nx, yy);
}
double orc_spline_eval(const orc_state *s, const double *coef, double xx, double yy)
{
  return eval_xy(s, coef, xx, yy);
}
