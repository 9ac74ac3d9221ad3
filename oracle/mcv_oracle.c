/*
 * mcv_oracle.c -- CPU restatement of the MCV_SW FP64 time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only the
 * test suite, the smoke check and the benchmark's CPU-baseline leg load it.
 *
 * Every routine names the reference file:line it restates.  Expressions keep
 * the reference's operand order and association, and the file is compiled
 * with -ffp-contract=off, so each double is rounded where the reference
 * rounds it.  Where the reference (built with -fdefault-real-8) promotes
 * dble() to REAL(16), the restatement uses __float128 for that expression.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   mesh, ghost tables, tendency, boundary pipeline, RK4/RK3 trajectories and
 *   diagnostics are checked against vectors written by the reference build
 *   (oracle/_ref, reference modules compiled unmodified); the operator is also
 *   checked against closed forms (polynomial exactness, rest state, RK
 *   polynomial) so that a dropped term or swapped index fails on its own.
 */
#define _GNU_SOURCE /* sincos */
#include "mcv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* constants_mod.F90:12-20 */
#define ORC_RADIUS 6371220.0
#define ORC_OMEGA 7.292e-5
#define ORC_FILL (-9999999999999999.0)

typedef struct {
  int dof, h, n, npv, nh;
  size_t np; /* points per field, halo included, 6 panels */
} geom;

static geom mkgeom(int dof, int n) {
  geom g;
  g.dof = dof;
  g.h = dof - 1;
  g.n = n;
  g.npv = n * g.h + 1;
  g.nh = g.npv + 2 * g.h;
  g.np = (size_t)g.nh * g.nh * 6;
  return g;
}

/* Fortran (i, j, p) with i, j in [1-H, nPV+H], p in [1, 6]. */
static inline size_t ix(const geom *g, int i, int j, int p) {
  return ((size_t)(p - 1) * g->nh + (size_t)(j - 1 + g->h)) * g->nh +
         (size_t)(i - 1 + g->h);
}

static double pi_ref(void) { return 2.0 * asin(1.0); } /* constants_mod.F90:13 */

static int valid(int dof, int n) { return (dof == 3 || dof == 4) && n >= 2; }

/* ------------------------------------------------------------------------ */
/* Geometry: projection_mod.F90 and mesh_mod.F90                            */
/* ------------------------------------------------------------------------ */

/* projection_mod.F90:8-38 */
static void plane2sphere(double x, double y, int p, double *lam, double *th) {
  const double pi = pi_ref();
  if (p <= 4) {
    /* lambda = x + dble(iPatch-1)*pi/2. is evaluated in REAL(16) */
    __float128 q = (__float128)x + ((__float128)(p - 1) * (__float128)pi) / (__float128)2.0;
    *lam = (double)q;
    *th = atan2(tan(y) * cos(x), 1.0);
  } else {
    double a = tan(x), b = tan(y);
    if (p == 5) {
      *lam = atan2(a, -b);
      *th = atan2(1.0, sqrt(a * a + b * b));
    } else {
      *lam = atan2(a, b);
      *th = -atan2(1.0, sqrt(a * a + b * b));
    }
  }
}

/* projection_mod.F90:40-58, panel 5 only (the ghost tables use panel 5). */
static void sphere2plane_p5(double lam, double th, double *x, double *y) {
  *x = atan((1.0 * sin(lam)) / tan(th));
  *y = atan(-cos(lam) / tan(th));
}

/* projection_mod.F90:150-170 */
static void metric_G(double x, double y, double *G) {
  double tx = tan(x), ty = tan(y), cx = cos(x), cy = cos(y);
  double rho = sqrt(1.0 + tx * tx + ty * ty);
  double m11 = 1.0 + tx * tx, m12 = -(tx * ty), m22 = 1.0 + ty * ty;
  double s = (ORC_RADIUS * ORC_RADIUS) / (pow(rho, 4.0) * (cx * cx) * (cy * cy));
  G[0] = s * m11;
  G[1] = s * m12;
  G[2] = s * m12;
  G[3] = s * m22;
}

/* projection_mod.F90:172-192 */
static void metric_IG(double x, double y, double *IG) {
  double tx = tan(x), ty = tan(y), cx = cos(x), cy = cos(y);
  double rho = sqrt(1.0 + tx * tx + ty * ty);
  double m11 = 1.0 + ty * ty, m12 = tx * ty, m22 = 1.0 + tx * tx;
  double s = (rho * rho * (cx * cx) * (cy * cy)) / (ORC_RADIUS * ORC_RADIUS);
  IG[0] = s * m11;
  IG[1] = s * m12;
  IG[2] = s * m12;
  IG[3] = s * m22;
}

/* projection_mod.F90:194-211 */
static double jacobian(double x, double y) {
  double tx = tan(x), ty = tan(y), cx = cos(x), cy = cos(y);
  double rho = sqrt(1.0 + tx * tx + ty * ty);
  double jab = 1.0 / ((cx * cx) * (cy * cy) * (rho * rho * rho));
  return jab * (ORC_RADIUS * ORC_RADIUS);
}

/* alambda = lambda - dble(iPatch-1)*pi/2. (REAL(16)), rounded to REAL(8) */
static double panel_lambda(double lam, int p) {
  const double pi = pi_ref();
  __float128 q = (__float128)lam - ((__float128)(p - 1) * (__float128)pi) / (__float128)2.0;
  return (double)q;
}

/* projection_mod.F90:213-267; M = (m11, m21, m12, m22) */
static void basis_IA(double lam, double th, int p, double *M) {
  double al = (p <= 4) ? panel_lambda(lam, p) : lam;
  double a, b, c, d, t;
  sincos(al, &a, &b); /* the reference build evaluates each sin/cos pair with sincos */
  sincos(th, &c, &d);
  double m11, m12, m21, m22;
  if (p <= 4) {
    t = d * d * b * b + c * c;
    m11 = 1.0 / d;
    m12 = 0.0;
    m21 = a * c / t;
    m22 = b / t;
  } else if (p == 5) {
    t = c + a * a * d * d / c;
    m11 = b / t;
    m12 = -(a / c / t);
    t = c + b * b * d * d / c;
    m21 = a / t;
    m22 = b / c / t;
  } else {
    t = c + a * a * d * d / c;
    m11 = -(b / t);
    m12 = a / c / t;
    t = c + b * b * d * d / c;
    m21 = a / t;
    m22 = b / c / t;
  }
  M[0] = m11 / ORC_RADIUS;
  M[1] = m21 / ORC_RADIUS;
  M[2] = m12 / ORC_RADIUS;
  M[3] = m22 / ORC_RADIUS;
}

/* projection_mod.F90:269-322 */
static void basis_A(double lam, double th, int p, double *M) {
  double al = (p <= 4) ? panel_lambda(lam, p) : lam;
  double a, b, c, d, t;
  sincos(al, &a, &b);
  sincos(th, &c, &d);
  double m11, m12, m21, m22;
  if (p <= 4) {
    m11 = d;
    m12 = 0.0;
    m21 = -(c * d * a / b);
    m22 = b * d * d + c * c / b;
  } else if (p == 5) {
    t = 1.0 + a * a * d * d / c / c;
    m11 = b * c * t;
    m21 = -(c * c * a * t);
    t = 1.0 + b * b * d * d / c / c;
    m12 = a * c * t;
    m22 = b * c * c * t;
  } else {
    t = 1.0 + a * a * d * d / c / c;
    m11 = -(b * c * t);
    m21 = c * c * a * t;
    t = 1.0 + b * b * d * d / c / c;
    m12 = a * c * t;
    m22 = b * c * c * t;
  }
  M[0] = m11 * ORC_RADIUS;
  M[1] = m21 * ORC_RADIUS;
  M[2] = m12 * ORC_RADIUS;
  M[3] = m22 * ORC_RADIUS;
}

/* mesh_mod.F90:160 calls sin alone; keep the compiler from fusing it with
 * the cos of the same argument above into a sincos call. */
static __attribute__((noinline)) double plain_sin(double x) { return sin(x); }

static void grid_spacing(int dof, int n, double *dx, double *xmin) {
  const double d2r = pi_ref() / 180.0;          /* constants_mod.F90:16 */
  *dx = (90.0 / (double)n) * d2r;               /* parameters_mod.F90:200 */
  *xmin = -45.0 * d2r;                          /* parameters_mod.F90:202 */
}

/* mesh_mod.F90:132-163 */
int orc_mesh_build(int dof, int n, double *xP, double *yP, double *lonP,
                   double *latP, double *sqrtG, double *G, double *IG,
                   double *A, double *IA, double *f) {
  if (!valid(dof, n)) return -1;
  geom g = mkgeom(dof, n);
  double dx, xmin;
  grid_spacing(dof, n, &dx, &xmin);
  for (int p = 1; p <= 6; ++p)
    for (int j = 1 - g.h; j <= g.npv + g.h; ++j)
      for (int i = 1 - g.h; i <= g.npv + g.h; ++i) {
        size_t k = ix(&g, i, j, p);
        double x = (double)(i - 1) * dx / (double)(dof - 1) + xmin;
        double y = (double)(j - 1) * dx / (double)(dof - 1) + xmin;
        double lam, th;
        plane2sphere(x, y, p, &lam, &th);
        xP[k] = x;
        yP[k] = y;
        lonP[k] = lam;
        latP[k] = th;
        metric_G(x, y, G + 4 * k);
        metric_IG(x, y, IG + 4 * k);
        basis_A(lam, th, p, A + 4 * k);
        basis_IA(lam, th, p, IA + 4 * k);
        sqrtG[k] = jacobian(x, y);
        f[k] = 2.0 * ORC_OMEGA * plain_sin(th);
      }
  return 0;
}

/* ghost_mod.f90:21-79 */
int orc_ghost_build(int dof, int n, const double *xP, const double *yP,
                    double *coef, int32_t *iref) {
  if (!valid(dof, n)) return -1;
  geom g = mkgeom(dof, n);
  double *xraw = malloc(sizeof(double) * (size_t)(n + 1));
  double *X = malloc(sizeof(double) * (size_t)g.npv * g.h);
  if (!xraw || !X) { free(xraw); free(X); return -2; }
  for (int k = 1; k <= g.h; ++k)
    for (int s = 1; s <= g.npv; ++s) {
      size_t m = ix(&g, s, g.npv + k, 1);
      double lam, th, y;
      plane2sphere(xP[m], yP[m], 1, &lam, &th);
      sphere2plane_p5(lam, th, &X[(size_t)(k - 1) * g.npv + (s - 1)], &y);
    }
  for (int c = 1; c <= n + 1; ++c) xraw[c - 1] = xP[ix(&g, (c - 1) * g.h + 1, 1, 1)];
  int rc = 0;
  for (int k = 1; k <= g.h && rc == 0; ++k) {
    int r = 1; /* 1-based index into xraw, reset per halo row */
    for (int s = 1; s <= g.npv; ++s) {
      double Xv = X[(size_t)(k - 1) * g.npv + (s - 1)];
      while (r <= n + 1 && Xv > xraw[r - 1]) ++r;
      if (r < 2 || r > n + 1 || !(Xv >= xraw[r - 2] && Xv <= xraw[r - 1])) { rc = -3; break; }
      coef[(size_t)(k - 1) * g.npv + (s - 1)] = Xv - xraw[r - 2];
      iref[(size_t)(k - 1) * g.npv + (s - 1)] = r - 1;
    }
  }
  free(xraw);
  free(X);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Spatial operator: spatial_operators_mod.F90:12-280                       */
/* ------------------------------------------------------------------------ */

/* spatial_operators_mod.F90:204-218 (polyfitL), pv[0..DOF-1] */
static inline double fitL(int dof, const double *pv, double dh) {
  if (dof == 3) return -((pv[2] - 4.0 * pv[1] + 3.0 * pv[0]) / dh);
  return (-(11.0 * pv[0]) + 18.0 * pv[1] - 9.0 * pv[2] + 2.0 * pv[3]) / (2.0 * dh);
}

/* spatial_operators_mod.F90:220-234 (polyfitR) */
static inline double fitR(int dof, const double *pv, double dh) {
  if (dof == 3) return (pv[0] - 4.0 * pv[1] + 3.0 * pv[2]) / dh;
  return (11.0 * pv[3] - 18.0 * pv[2] + 9.0 * pv[1] - 2.0 * pv[0]) / (2.0 * dh);
}

/* spatial_operators_mod.F90:236-246 */
static inline double riemann(double fxL, double fxR, double qxL, double qxR, double lam) {
  return 0.5 * ((fxL + fxR) - lam * (qxR - qxL));
}

/* spatial_operators_mod.F90:248-280 */
static inline double eigen(double c, double phi, double g) {
  double l1 = c - sqrt(g * phi), l2 = c + sqrt(g * phi);
  double a1 = fabs(l1), a2 = fabs(l2);
  return a1 > a2 ? a1 : a2;
}

/* spatial_operators_mod.F90:136-202.  f, q: 3*DOF-2 window; d: DOF values. */
static void tend_window(int dof, double dx, const double *f, const double *q,
                        double eL, double eR, double *d) {
  const int h = dof - 1;
  const double *f1 = f, *f2 = f + h, *f3 = f + 2 * h;
  const double *q1 = q, *q2 = q + h, *q3 = q + 2 * h;
  double fxL1 = fitR(dof, f1, dx), fxL2 = fitR(dof, f2, dx);
  double fxR1 = fitL(dof, f2, dx), fxR2 = fitL(dof, f3, dx);
  double qxL1 = fitR(dof, q1, dx), qxL2 = fitR(dof, q2, dx);
  double qxR1 = fitL(dof, q2, dx), qxR2 = fitL(dof, q3, dx);
  d[0] = riemann(fxL1, fxR1, qxL1, qxR1, eL);
  d[h] = riemann(fxL2, fxR2, qxL2, qxR2, eR);
  if (dof == 3) {
    double via = (f2[2] - f2[0]) / dx;
    d[1] = 1.5 * via - 0.25 * (d[0] + d[2]);
  } else {
    double via = (f2[3] - f2[0]) / dx;
    double fxx = 4.5 * (f2[0] - f2[1] - f2[2] + f2[3]) / (dx * dx);
    d[1] = 4.0 / 3.0 * via - (4.0 * d[0] + 5.0 * d[3]) / 27.0 - 4.0 / 27.0 * dx * fxx;
    d[2] = 4.0 / 3.0 * via - (5.0 * d[0] + 4.0 * d[3]) / 27.0 + 4.0 / 27.0 * dx * fxx;
  }
}

/* x/|x| with 0 -> 1 (spatial_operators_mod.F90:83-85) */
static inline double upstream(double c) {
  if (c == 0.0) c = 1.0;
  return c / fabs(c);
}

/* interior index into (nPV, nPV, 6) tendency arrays */
static inline size_t tix(const geom *g, int i, int j, int p) {
  return ((size_t)(p - 1) * g->npv + (size_t)(j - 1)) * g->npv + (size_t)(i - 1);
}

int orc_tendency(int dof, int n, double dx, const double *sqrtG,
                 const double *IG, const double *f, const double *phi_s,
                 const double *phi, const double *phiG, const double *u,
                 const double *v, const double *cu, const double *cv,
                 double *t_phiG, double *t_u, double *t_v) {
  if (!valid(dof, n)) return -1;
  geom g = mkgeom(dof, n);
  const int h = g.h, w = 3 * dof - 2;
  const size_t ni = (size_t)g.npv * g.npv * 6;
  double *E = malloc(sizeof(double) * g.np), *pgu = malloc(sizeof(double) * g.np),
         *pgv = malloc(sizeof(double) * g.np);
  double *fx = malloc(sizeof(double) * ni), *fy = malloc(sizeof(double) * ni),
         *dvx = malloc(sizeof(double) * ni), *dvy = malloc(sizeof(double) * ni),
         *vdx = malloc(sizeof(double) * ni), *udy = malloc(sizeof(double) * ni);
  if (!E || !pgu || !pgv || !fx || !fy || !dvx || !dvy || !vdx || !udy) {
    free(E); free(pgu); free(pgv); free(fx); free(fy); free(dvx); free(dvy); free(vdx); free(udy);
    return -2;
  }
  /* K1: spatial_operators_mod.F90:58-61 */
  for (size_t k = 0; k < g.np; ++k) {
    pgu[k] = phiG[k] * cu[k];
    pgv[k] = phiG[k] * cv[k];
    double K = 0.5 * (cu[k] * u[k] + cv[k] * v[k]);
    E[k] = phi[k] + phi_s[k] + K;
  }
  double wE[10], wa[10], wb[10], wc[10], d[4];
  /* K2: x sweep, spatial_operators_mod.F90:64-93 */
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= g.npv; ++j)
      for (int c = 1; c <= n; ++c) {
        int Pl = (c - 1) * h + 1, Pr = Pl + h, Ps = Pl - h;
        for (int t = 0; t < w; ++t) {
          size_t k = ix(&g, Ps + t, j, p);
          wE[t] = E[k]; wa[t] = u[k]; wb[t] = pgu[k]; wc[t] = phiG[k];
        }
        size_t kl = ix(&g, Pl, j, p), kr = ix(&g, Pr, j, p);
        double lL = eigen(cu[kl], phi[kl], IG[4 * kl]);
        double lR = eigen(cu[kr], phi[kr], IG[4 * kr]);
        double uL = upstream(cu[kl]), uR = upstream(cu[kr]);
        tend_window(dof, dx, wE, wa, lL, lR, d);
        for (int t = 0; t < dof; ++t) fx[tix(&g, Pl + t, j, p)] = d[t];
        tend_window(dof, dx, wb, wc, lL, lR, d);
        for (int t = 0; t < dof; ++t) dvx[tix(&g, Pl + t, j, p)] = d[t];
        for (int t = 0; t < w; ++t) wa[t] = v[ix(&g, Ps + t, j, p)];
        tend_window(dof, dx, wa, wa, uL, uR, d);
        for (int t = 0; t < dof; ++t) vdx[tix(&g, Pl + t, j, p)] = d[t];
      }
  /* K3: y sweep, spatial_operators_mod.F90:96-125 */
  for (int p = 1; p <= 6; ++p)
    for (int i = 1; i <= g.npv; ++i)
      for (int c = 1; c <= n; ++c) {
        int Pl = (c - 1) * h + 1, Pr = Pl + h, Ps = Pl - h;
        for (int t = 0; t < w; ++t) {
          size_t k = ix(&g, i, Ps + t, p);
          wE[t] = E[k]; wa[t] = v[k]; wb[t] = pgv[k]; wc[t] = phiG[k];
        }
        size_t kl = ix(&g, i, Pl, p), kr = ix(&g, i, Pr, p);
        double lL = eigen(cv[kl], phi[kl], IG[4 * kl + 3]);
        double lR = eigen(cv[kr], phi[kr], IG[4 * kr + 3]);
        double uL = upstream(cv[kl]), uR = upstream(cv[kr]);
        tend_window(dof, dx, wE, wa, lL, lR, d);
        for (int t = 0; t < dof; ++t) fy[tix(&g, i, Pl + t, p)] = d[t];
        tend_window(dof, dx, wb, wc, lL, lR, d);
        for (int t = 0; t < dof; ++t) dvy[tix(&g, i, Pl + t, p)] = d[t];
        for (int t = 0; t < w; ++t) wa[t] = u[ix(&g, i, Ps + t, p)];
        tend_window(dof, dx, wa, wa, uL, uR, d);
        for (int t = 0; t < dof; ++t) udy[tix(&g, i, Pl + t, p)] = d[t];
      }
  /* K4: spatial_operators_mod.F90:127-131 */
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= g.npv; ++j)
      for (int i = 1; i <= g.npv; ++i) {
        size_t k = ix(&g, i, j, p), m = tix(&g, i, j, p);
        double vort = (vdx[m] - udy[m]) / sqrtG[k];
        t_phiG[m] = -dvx[m] - dvy[m];
        t_u[m] = -fx[m] + sqrtG[k] * (vort + f[k]) * cv[k];
        t_v[m] = -fy[m] - sqrtG[k] * (vort + f[k]) * cu[k];
      }
  free(E); free(pgu); free(pgv); free(fx); free(fy); free(dvx); free(dvy); free(vdx); free(udy);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Boundary pipeline: time_scheme_mod.F90:125-142                           */
/* ------------------------------------------------------------------------ */

/* projection_mod.F90:79-96 with the product formed first: (A*IG) */
static inline void p2sp(const double *A, const double *IG, double cu1, double cu2,
                        double *s1, double *s2) {
  double m11 = A[0] * IG[0] + A[2] * IG[1];
  double m21 = A[1] * IG[0] + A[3] * IG[1];
  double m12 = A[0] * IG[2] + A[2] * IG[3];
  double m22 = A[1] * IG[2] + A[3] * IG[3];
  *s1 = m11 * cu1 + m12 * cu2;
  *s2 = m21 * cu1 + m22 * cu2;
}

/* projection_mod.F90:60-77 with the product formed first: (G*IA) */
static inline void sp2p(const double *G, const double *IA, double s1, double s2,
                        double *c1, double *c2) {
  double m11 = G[0] * IA[0] + G[2] * IA[1];
  double m21 = G[1] * IA[0] + G[3] * IA[1];
  double m12 = G[0] * IA[2] + G[2] * IA[3];
  double m22 = G[1] * IA[2] + G[3] * IA[3];
  *c1 = m11 * s1 + m12 * s2;
  *c2 = m21 * s1 + m22 * s2;
}

/* Halo source of CubedSphereFillHalo (ghost_mod.f90:88-152).
 * side: 0 W (i=1-k), 1 E (i=nPV+k), 2 S (j=1-k), 3 N (j=nPV+k).
 * For panel p, side, depth k and along-edge index s it returns the interior
 * point (q, i, j) that is copied into the halo. */
typedef struct { int q, col, far, rev; } halo_src;
static const halo_src HALO[6][4] = {
    /* P1 */ {{4, 1, 1, 0}, {2, 1, 0, 0}, {6, 0, 1, 0}, {5, 0, 0, 0}},
    /* P2 */ {{1, 1, 1, 0}, {3, 1, 0, 0}, {6, 1, 1, 1}, {5, 1, 1, 0}},
    /* P3 */ {{2, 1, 1, 0}, {4, 1, 0, 0}, {6, 0, 0, 1}, {5, 0, 1, 1}},
    /* P4 */ {{3, 1, 1, 0}, {1, 1, 0, 0}, {6, 1, 0, 0}, {5, 1, 0, 1}},
    /* P5 */ {{4, 0, 1, 1}, {2, 0, 1, 0}, {1, 0, 1, 0}, {3, 0, 1, 1}},
    /* P6 */ {{4, 0, 0, 0}, {2, 0, 0, 1}, {3, 0, 0, 1}, {1, 0, 0, 0}},
};

static void halo_source(const geom *g, int p, int side, int k, int s, int *q, int *i, int *j) {
  const halo_src *hs = &HALO[p - 1][side];
  int nrm = hs->far ? g->npv - k : k + 1;
  int al = hs->rev ? g->npv + 1 - s : s;
  *q = hs->q;
  if (hs->col) { *i = nrm; *j = al; } else { *i = al; *j = nrm; }
}

/* halo point (i, j) of panel side at depth k, along index s */
static void halo_point(const geom *g, int side, int k, int s, int *i, int *j) {
  switch (side) {
    case 0: *i = 1 - k; *j = s; break;
    case 1: *i = g->npv + k; *j = s; break;
    case 2: *i = s; *j = 1 - k; break;
    default: *i = s; *j = g->npv + k; break;
  }
}

/* ghost_mod.f90:183-227 (linear_interp, degree DOF-1 Lagrange) */
static double interp(int dof, double dx, const double *q, double c) {
  if (dof == 3) {
    return q[0] - (3.0 * q[0] - 4.0 * q[1] + q[2]) * c / dx +
           2.0 * (q[0] - 2.0 * q[1] + q[2]) * (c * c) / (dx * dx);
  }
  double dx3 = dx * dx * dx, c3 = c * c * c; /* x**3. is evaluated as an integer power */
  return (2.0 * dx3 * q[0] + (dx * dx) * (-(11.0 * q[0]) + 18.0 * q[1] - 9.0 * q[2] + 2.0 * q[3]) * c +
          9.0 * dx * (2.0 * q[0] - 5.0 * q[1] + 4.0 * q[2] - q[3]) * (c * c) +
          9.0 * (-q[0] + 3.0 * q[1] - 3.0 * q[2] + q[3]) * c3) /
         (2.0 * dx3);
}

/* ghost_mod.f90:154-181 (CubedSphereFillGhost) for one field, in place. */
static void fill_ghost_field(const geom *g, double dx, const double *coef,
                             const int32_t *iref, double *F, double *T, double *line) {
  const int h = g->h, npv = g->npv;
  for (size_t k = 0; k < g->np; ++k) T[k] = ORC_FILL;
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= npv; ++j)
      for (int i = 1; i <= npv; ++i) T[ix(g, i, j, p)] = F[ix(g, i, j, p)];
  /* CubedSphereFillHalo copies the neighbour line into F's halo; the
   * interpolation then reads that copied line.  We gather it into `line`. */
  for (int p = 1; p <= 6; ++p)
    for (int side = 0; side < 4; ++side)
      for (int k = 1; k <= h; ++k) {
        for (int s = 1; s <= npv; ++s) {
          int q, si, sj;
          halo_source(g, p, side, k, s, &q, &si, &sj);
          line[s - 1] = F[ix(g, si, sj, q)];
        }
        for (int s = 1; s <= npv; ++s) {
          int c = iref[(size_t)(k - 1) * npv + (s - 1)];
          double cf = coef[(size_t)(k - 1) * npv + (s - 1)];
          int hi, hj;
          halo_point(g, side, k, s, &hi, &hj);
          T[ix(g, hi, hj, p)] = interp(g->dof, dx, line + (c - 1) * h, cf);
        }
      }
  memcpy(F, T, sizeof(double) * g->np);
}

/* One unified cube edge of spatial_operators_mod.F90:298-345:
 * (panel a, side a) and (panel b, side b), b traversed reversed if rev. */
typedef struct { int pa, sa, pb, sb, rev; } edge_pair;
static const edge_pair EDGES[12] = {
    {1, 0, 4, 1, 0}, {2, 0, 1, 1, 0}, {3, 0, 2, 1, 0}, {4, 0, 3, 1, 0},
    {1, 3, 5, 2, 0}, {1, 2, 6, 3, 0}, {2, 3, 5, 1, 0}, {2, 2, 6, 1, 1},
    {3, 3, 5, 3, 1}, {3, 2, 6, 2, 1}, {4, 3, 5, 0, 1}, {4, 2, 6, 0, 0},
};

/* Vertices of spatial_operators_mod.F90:347-386, in the reference's
 * summation order: (panel, i-corner, j-corner) with 0 = 1, 1 = nPV. */
typedef struct { int p, ci, cj; } vtx_pt;
static const vtx_pt VERTS[8][3] = {
    {{1, 0, 1}, {4, 1, 1}, {5, 0, 0}}, {{1, 1, 1}, {2, 0, 1}, {5, 1, 0}},
    {{1, 0, 0}, {4, 1, 0}, {6, 0, 1}}, {{1, 1, 0}, {2, 0, 0}, {6, 1, 1}},
    {{3, 0, 1}, {2, 1, 1}, {5, 1, 1}}, {{3, 1, 1}, {4, 0, 1}, {5, 0, 1}},
    {{2, 1, 0}, {3, 0, 0}, {6, 1, 0}}, {{3, 1, 0}, {4, 0, 0}, {6, 0, 0}},
};

static void edge_point(const geom *g, int side, int s, int *i, int *j) {
  switch (side) {
    case 0: *i = 1; *j = s; break;
    case 1: *i = g->npv; *j = s; break;
    case 2: *i = s; *j = 1; break;
    default: *i = s; *j = g->npv; break;
  }
}

/* spatial_operators_mod.F90:291-388 (unify_bdy_field), in place. */
static void unify_field(const geom *g, double *F, double *raw) {
  memcpy(raw, F, sizeof(double) * g->np);
  for (int e = 0; e < 12; ++e) {
    const edge_pair *ep = &EDGES[e];
    for (int s = 1; s <= g->npv; ++s) {
      int ia, ja, ib, jb;
      edge_point(g, ep->sa, s, &ia, &ja);
      edge_point(g, ep->sb, ep->rev ? g->npv + 1 - s : s, &ib, &jb);
      double val = 0.5 * (raw[ix(g, ia, ja, ep->pa)] + raw[ix(g, ib, jb, ep->pb)]);
      F[ix(g, ia, ja, ep->pa)] = val;
      F[ix(g, ib, jb, ep->pb)] = val;
    }
  }
  for (int vtx = 0; vtx < 8; ++vtx) {
    size_t k[3];
    for (int t = 0; t < 3; ++t) {
      const vtx_pt *vp = &VERTS[vtx][t];
      k[t] = ix(g, vp->ci ? g->npv : 1, vp->cj ? g->npv : 1, vp->p);
    }
    double val = (raw[k[0]] + raw[k[1]] + raw[k[2]]) / 3.0;
    for (int t = 0; t < 3; ++t) F[k[t]] = val;
  }
}

typedef struct {
  const double *sqrtG, *G, *IG, *A, *IA, *f, *phi_s, *coef;
  const int32_t *iref;
  double dx;
} mesh_in;

typedef struct { double *phi, *phiG, *u, *v, *cu, *cv, *zon, *mer; } st;

static void boundary(const geom *g, const mesh_in *m, st *s, double *w1, double *w2) {
  const int h = g->h, npv = g->npv;
  /* convert_ghost_wind_P2SP: spatial_operators_mod.F90:408-441 */
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= npv; ++j)
      for (int i = 1; i <= npv; ++i) {
        int band = (i <= 1 + h) || (i >= npv - h) || (j <= 1 + h) || (j >= npv - h);
        if (!band) continue;
        size_t k = ix(g, i, j, p);
        p2sp(m->A + 4 * k, m->IG + 4 * k, s->u[k], s->v[k], &s->zon[k], &s->mer[k]);
      }
  /* fill_ghost: ghost_mod.f90:230-237 */
  fill_ghost_field(g, m->dx, m->coef, m->iref, s->phi, w1, w2);
  fill_ghost_field(g, m->dx, m->coef, m->iref, s->zon, w1, w2);
  fill_ghost_field(g, m->dx, m->coef, m->iref, s->mer, w1, w2);
  /* unify_bdy_stat: spatial_operators_mod.F90:282-289 */
  unify_field(g, s->phi, w1);
  unify_field(g, s->zon, w1);
  unify_field(g, s->mer, w1);
  /* convert_ghost_wind_SP2P: spatial_operators_mod.F90:459-492 (halo strips
   * plus the edge lines, corners excluded) */
  for (int p = 1; p <= 6; ++p)
    for (int j = 1 - h; j <= npv + h; ++j)
      for (int i = 1 - h; i <= npv + h; ++i) {
        int in_i = (i >= 1 && i <= npv), in_j = (j >= 1 && j <= npv);
        int strip = (in_j && (i <= 1 || i >= npv)) || (in_i && (j <= 1 || j >= npv));
        if (!strip) continue;
        size_t k = ix(g, i, j, p);
        sp2p(m->G + 4 * k, m->IA + 4 * k, s->zon[k], s->mer[k], &s->u[k], &s->v[k]);
      }
  /* convert_wind_cov2contrav: spatial_operators_mod.F90:494-508 */
  for (size_t k = 0; k < g->np; ++k) {
    const double *IG = m->IG + 4 * k;
    s->cu[k] = IG[0] * s->u[k] + IG[2] * s->v[k];
    s->cv[k] = IG[1] * s->u[k] + IG[3] * s->v[k];
  }
}

int orc_boundary(int dof, int n, double dx, const double *G, const double *IG,
                 const double *A, const double *IA, const double *coef,
                 const int32_t *iref, double *phi, double *u, double *v,
                 double *contraU, double *contraV, double *zonal, double *merid) {
  if (!valid(dof, n)) return -1;
  geom g = mkgeom(dof, n);
  mesh_in m = {0};
  m.G = G; m.IG = IG; m.A = A; m.IA = IA; m.coef = coef; m.iref = iref; m.dx = dx;
  st s = {phi, NULL, u, v, contraU, contraV, zonal, merid};
  double *w1 = malloc(sizeof(double) * g.np), *w2 = malloc(sizeof(double) * g.npv);
  if (!w1 || !w2) { free(w1); free(w2); return -2; }
  boundary(&g, &m, &s, w1, w2);
  free(w1);
  free(w2);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Time stepping: time_scheme_mod.F90:13-123, stat_mod.F90:47-73           */
/* ------------------------------------------------------------------------ */

typedef struct { double *phiG, *u, *v; } td;

static void copy_state(const geom *g, st *dst, const st *src) {
  size_t b = sizeof(double) * g->np;
  memcpy(dst->phi, src->phi, b); memcpy(dst->phiG, src->phiG, b);
  memcpy(dst->u, src->u, b); memcpy(dst->v, src->v, b);
  memcpy(dst->cu, src->cu, b); memcpy(dst->cv, src->cv, b);
  memcpy(dst->zon, src->zon, b); memcpy(dst->mer, src->mer, b);
}

static void op(const geom *g, const mesh_in *m, const st *s, td *t) {
  orc_tendency(g->dof, g->n, m->dx, m->sqrtG, m->IG, m->f, m->phi_s, s->phi, s->phiG,
               s->u, s->v, s->cu, s->cv, t->phiG, t->u, t->v);
}

/* time_scheme_mod.F90:81-85 */
static void finish(const geom *g, const mesh_in *m, st *s, double *w1, double *w2) {
  for (size_t k = 0; k < g->np; ++k) s->phi[k] = s->phiG[k] / m->sqrtG[k];
  boundary(g, m, s, w1, w2);
  for (size_t k = 0; k < g->np; ++k) s->phiG[k] = s->phi[k] * m->sqrtG[k];
}

/* kind 0: Yn + inc*k (update_stat); 1: 0.75 Yn + 0.25 Y1 + 0.25 dt k;
 * 2: Yn/3 + 2/3 Y2 + 2/3 dt k (update_stat_RK3_TVD_1/_2). */
static void update(const geom *g, const mesh_in *m, int kind, double inc, double dt,
                   st *sn, const st *sb, const st *s1, const td *t, double *w1, double *w2) {
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= g->npv; ++j)
      for (int i = 1; i <= g->npv; ++i) {
        size_t k = ix(g, i, j, p), mm = tix(g, i, j, p);
        const double *b[3] = {sb->phiG, sb->u, sb->v};
        const double *y[3] = {s1 ? s1->phiG : NULL, s1 ? s1->u : NULL, s1 ? s1->v : NULL};
        const double *tt[3] = {t->phiG, t->u, t->v};
        double *o[3] = {sn->phiG, sn->u, sn->v};
        for (int fl = 0; fl < 3; ++fl) {
          if (kind == 0)
            o[fl][k] = b[fl][k] + inc * tt[fl][mm];
          else if (kind == 1)
            o[fl][k] = 0.75 * b[fl][k] + 0.25 * y[fl][k] + 0.25 * dt * tt[fl][mm];
          else
            o[fl][k] = b[fl][k] / 3.0 + 2.0 / 3.0 * y[fl][k] + 2.0 / 3.0 * dt * tt[fl][mm];
        }
      }
  finish(g, m, sn, w1, w2);
}

int orc_run(int dof, int n, int scheme, double dt, double dx,
            const double *sqrtG, const double *G, const double *IG,
            const double *A, const double *IA, const double *f,
            const double *phi_s, const double *coef, const int32_t *iref,
            double *phi, double *phiG, double *u, double *v,
            double *contraU, double *contraV, double *zonal, double *merid,
            int nsteps) {
  if (!valid(dof, n) || (scheme != 3 && scheme != 4) || nsteps < 0) return -1;
  geom g = mkgeom(dof, n);
  mesh_in m = {sqrtG, G, IG, A, IA, f, phi_s, coef, iref, dx};
  const size_t ni = (size_t)g.npv * g.npv * 6;
  /* levels: 0 = new, 1..4 = stat(-1..-4); tends 0 = new, 1..4 = tend(-1..-4) */
  st lv[5];
  td tdv[5];
  double *pool = calloc(5 * 8 * g.np + 5 * 3 * ni + g.np + g.npv, sizeof(double));
  if (!pool) return -2;
  double *cur = pool;
  for (int l = 0; l < 5; ++l) {
    double **fields[8] = {&lv[l].phi, &lv[l].phiG, &lv[l].u, &lv[l].v,
                          &lv[l].cu, &lv[l].cv, &lv[l].zon, &lv[l].mer};
    for (int fl = 0; fl < 8; ++fl) { *fields[fl] = cur; cur += g.np; }
  }
  for (int l = 0; l < 5; ++l) {
    tdv[l].phiG = cur; cur += ni;
    tdv[l].u = cur; cur += ni;
    tdv[l].v = cur; cur += ni;
  }
  double *w1 = cur; cur += g.np;
  double *w2 = cur;
  st old = {phi, phiG, u, v, contraU, contraV, zonal, merid};
  st *nw = &lv[0];
  for (int it = 0; it < nsteps; ++it) {
    copy_state(&g, &lv[1], &old);
    if (scheme == 4) {
      op(&g, &m, &lv[1], &tdv[1]);
      update(&g, &m, 0, 0.5 * dt, dt, &lv[2], &lv[1], NULL, &tdv[1], w1, w2);
      op(&g, &m, &lv[2], &tdv[2]);
      update(&g, &m, 0, 0.5 * dt, dt, &lv[3], &lv[1], NULL, &tdv[2], w1, w2);
      op(&g, &m, &lv[3], &tdv[3]);
      update(&g, &m, 0, dt, dt, &lv[4], &lv[1], NULL, &tdv[3], w1, w2);
      op(&g, &m, &lv[4], &tdv[4]);
      for (size_t k = 0; k < ni; ++k) {
        tdv[0].phiG[k] = (tdv[1].phiG[k] + 2.0 * tdv[2].phiG[k] + 2.0 * tdv[3].phiG[k] + tdv[4].phiG[k]) / 6.0;
        tdv[0].u[k] = (tdv[1].u[k] + 2.0 * tdv[2].u[k] + 2.0 * tdv[3].u[k] + tdv[4].u[k]) / 6.0;
        tdv[0].v[k] = (tdv[1].v[k] + 2.0 * tdv[2].v[k] + 2.0 * tdv[3].v[k] + tdv[4].v[k]) / 6.0;
      }
      update(&g, &m, 0, dt, dt, nw, &old, NULL, &tdv[0], w1, w2);
    } else {
      op(&g, &m, &lv[1], &tdv[1]);
      update(&g, &m, 0, dt, dt, &lv[2], &lv[1], NULL, &tdv[1], w1, w2);
      op(&g, &m, &lv[2], &tdv[2]);
      update(&g, &m, 1, 0.0, dt, &lv[3], &lv[1], &lv[2], &tdv[2], w1, w2);
      op(&g, &m, &lv[3], &tdv[3]);
      update(&g, &m, 2, 0.0, dt, nw, &lv[1], &lv[3], &tdv[3], w1, w2);
    }
    copy_state(&g, &old, nw); /* switch_stat, MCV_SW.F90:76 */
  }
  free(pool);
  return 0;
}

/* diag_mod.F90:10-67 */
int orc_diag(int dof, int n, const double *sqrtG, const double *phi_s,
             const double *phi, const double *phiG, const double *u,
             const double *v, const double *cu, const double *cv,
             double *mass, double *energy) {
  if (!valid(dof, n)) return -1;
  geom g = mkgeom(dof, n);
  double w1[4];
  if (dof == 3) { w1[0] = 1.0; w1[1] = 4.0; w1[2] = 1.0; }
  else { w1[0] = 1.0; w1[1] = 3.0; w1[2] = 3.0; w1[3] = 1.0; }
  double w[4][4], norm = dof == 3 ? 36.0 : 80.0;
  for (int a = 0; a < dof; ++a)
    for (int b = 0; b < dof; ++b) w[a][b] = (w1[a] * w1[b]) / norm; /* mesh_mod.F90:183-203 */
  double msum = 0.0, esum = 0.0;
  for (int p = 1; p <= 6; ++p)
    for (int cj = 1; cj <= n; ++cj)
      for (int ci = 1; ci <= n; ++ci) {
        double mc = 0.0, ec = 0.0;
        for (int b = 0; b < dof; ++b)
          for (int a = 0; a < dof; ++a) {
            size_t k = ix(&g, (ci - 1) * g.h + 1 + a, (cj - 1) * g.h + 1 + b, p);
            mc = mc + w[a][b] * phiG[k];
            double KE = 0.5 * phi[k] * (u[k] * cu[k] + v[k] * cv[k]);
            double PE = 0.5 * ((phi[k] + phi_s[k]) * (phi[k] + phi_s[k]));
            ec = ec + w[a][b] * sqrtG[k] * (KE + PE);
          }
        msum += mc;
        esum += ec;
      }
  *mass = msum;
  *energy = esum;
  return 0;
}
