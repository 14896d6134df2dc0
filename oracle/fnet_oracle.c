/*
 * fnet_oracle.c -- FP64 CPU restatement of Fortnet's BPNN hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only the tests, the
 * smoke check and the CPU-baseline leg of the benchmark may load it.  The
 * product path (fortnet_amd + libfnet_hip.so) never links or calls it.
 *
 * Every routine follows the reference's algorithm (file:line of
 * vanderhe/fortnet 0.7.4, prog/fortnet/...) in plain double precision and in the
 * reference's own operation order where that order decides a comparison (the
 * neighbour distance test).  It is pinned by the reference's HDF5 regression
 * fixtures (converted to .npz under tests/golden), by closed forms and by
 * finite differences; see the oracle tests.
 *
 * The ten activations (act_f/act_d, transfer.F90:54-342) are pinned by the reference's
 * scratch-training fixtures (one steepest-descent step with each transfer function).
 *
 * Layouts (row-major C == Fortran column-major with the indices reversed):
 *   coords[n][3]           Fortran coords(3, nAtom)
 *   lat[3][3]              lat[c][r] == Fortran latVecs(r, c) (row c = lattice vector c)
 *   G[n][nG]               Fortran vals(nG, nAtom)
 *   dG[n][n][nG][3]        Fortran valsPrime(3, nG, nAtom, nAtom): dG[k][i][a][c]
 *   ext[n][next]           Fortran extFeatures(nExt, nAtom)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846 /* lib_dftbp/constants.F90:15 */

/* mode flags */
#define ORC_REFSHAPE 1      /* rebuild the neighbour list per (atom, function) like acsf.F90:829-834 */
#define ORC_COMPAT_SCATTER 2 /* last-write-wins dense scatter of acsf.F90:918-933 */
#define ORC_COMPAT_SELFIMG 4 /* keep the centre's own images in species-resolved lists, acsf.F90:754-760 */
/* A planted mistake for the tests (never a reference behaviour): cos(theta) of the G4/G5 values
 * as dot * (1/r_j) * (1/r_k) without the 1e-13 regulariser of acsf.F90:1173-1174, and exactly 1
 * for j == k.  The tests show that the reference's element-wise tolerance detects it. */
#define ORC_UNREG_COS 8

typedef struct {
  int nG;
  const int* type;      /* 1..5 = G1..G5 */
  const double* rcut;   /* Bohr */
  const double* eta;
  const double* rs;
  const double* kappa;
  const double* lam;
  const double* xi;
  const int* atomid;    /* 0 = none, else 1-based column of ext */
  const int* z1;
  const int* z2;
} orc_gfuncs;

/* ------------------------------------------------------------------------- */
/* geometry helpers: simplealgebra.F90:41-62 (determinant33), :90-125 (invert33) */

static double det33(const double* lat) {
  /* o(r,c) = lat[c*3+r] */
#define O(r, c) lat[((c) - 1) * 3 + ((r) - 1)]
  double tmp = O(1, 1) * (O(2, 2) * O(3, 3) - O(3, 2) * O(2, 3));
  tmp = tmp - O(1, 2) * (O(2, 1) * O(3, 3) - O(3, 1) * O(2, 3));
  tmp = tmp + O(1, 3) * (O(2, 1) * O(3, 2) - O(3, 1) * O(2, 2));
  return tmp;
}

/* inv(r,c) stored as inv[c*3+r] (same convention as lat) */
static void inv33(const double* lat, double* inv) {
  double det = det33(lat);
#define I(r, c) inv[((c) - 1) * 3 + ((r) - 1)]
  I(1, 1) = -O(2, 3) * O(3, 2) + O(2, 2) * O(3, 3);
  I(2, 1) = O(2, 3) * O(3, 1) - O(2, 1) * O(3, 3);
  I(3, 1) = -O(2, 2) * O(3, 1) + O(2, 1) * O(3, 2);
  I(1, 2) = O(1, 3) * O(3, 2) - O(1, 2) * O(3, 3);
  I(2, 2) = -O(1, 3) * O(3, 1) + O(1, 1) * O(3, 3);
  I(3, 2) = O(1, 2) * O(3, 1) - O(1, 1) * O(3, 2);
  I(1, 3) = -O(1, 3) * O(2, 2) + O(1, 2) * O(2, 3);
  I(2, 3) = O(1, 3) * O(2, 1) - O(1, 1) * O(2, 3);
  I(3, 3) = -O(1, 2) * O(2, 1) + O(1, 1) * O(2, 2);
  for (int q = 0; q < 9; ++q) inv[q] = inv[q] / det;
#undef I
#undef O
}

/* Image ranges: latpointiter.F90:226-253 with neg/posExtension = 1 (dynneighlist.F90:223-224).
 * recVecs2p(:, ii) = row ii of inv(latVecs) (typegeometry.F90:181-182, fnetdata.F90:1155-1157). */
void orc_image_ranges(const double* lat, double rc, int* lo, int* hi) {
  double inv[9];
  inv33(lat, inv);
  for (int ii = 0; ii < 3; ++ii) {
    /* row ii of inv: inv(ii, c) = inv[c*3+ii] */
    double s = 0.0;
    for (int c = 0; c < 3; ++c) {
      double v = inv[c * 3 + ii];
      s = s + v * v;
    }
    int t = (int)floor(rc * sqrt(s));
    lo[ii] = -(t + 1);
    hi[ii] = t + 1;
  }
}

/* ------------------------------------------------------------------------- */
/* Neighbour iterator: dynneighlist.F90:182-320 and latpointiter.F90:121-222.
 * Emits entries in the reference's order: central cell j = i+1..n-1, 0..i-1, then
 * for every lattice point (third index fastest, origin skipped) j = i..n-1, 0..i-1.
 * Returns the number of entries; writes at most cap of them. */
long orc_neighbours(int n, const double* xyz, const double* lat, int periodic, double rc, int i,
                    long cap, int* jidx, int* imgidx, double* ncoord, double* dist2) {
  const double rc2 = rc * rc;
  const double* ci = xyz + 3 * i;
  long cnt = 0;
  for (int t = 1; t < n; ++t) {
    int j = (i + t) % n;
    double nc0 = xyz[3 * j + 0] + 0.0, nc1 = xyz[3 * j + 1] + 0.0, nc2 = xyz[3 * j + 2] + 0.0;
    double d0 = ci[0] - nc0, d1 = ci[1] - nc1, d2 = ci[2] - nc2;
    double s = d0 * d0;
    s = s + d1 * d1;
    s = s + d2 * d2;
    if (s <= rc2) {
      if (cnt < cap) {
        jidx[cnt] = j;
        if (imgidx) imgidx[cnt] = -1;
        ncoord[3 * cnt + 0] = nc0;
        ncoord[3 * cnt + 1] = nc1;
        ncoord[3 * cnt + 2] = nc2;
        dist2[cnt] = s;
      }
      ++cnt;
    }
  }
  if (!periodic) return cnt;
  int lo[3], hi[3];
  orc_image_ranges(lat, rc, lo, hi);
  int img = 0;
  for (int p1 = lo[0]; p1 <= hi[0]; ++p1)
    for (int p2 = lo[1]; p2 <= hi[1]; ++p2)
      for (int p3 = lo[2]; p3 <= hi[2]; ++p3) {
        if (p1 == 0 && p2 == 0 && p3 == 0) continue;
        double pv[3] = {(double)p1, (double)p2, (double)p3};
        double cell[3];
        for (int r = 0; r < 3; ++r) {
          double acc = 0.0;
          for (int c = 0; c < 3; ++c) acc = acc + lat[c * 3 + r] * pv[c];
          cell[r] = acc;
        }
        for (int t = 0; t < n; ++t) {
          int j = (i + t) % n;
          double nc0 = xyz[3 * j + 0] + cell[0];
          double nc1 = xyz[3 * j + 1] + cell[1];
          double nc2 = xyz[3 * j + 2] + cell[2];
          double d0 = ci[0] - nc0, d1 = ci[1] - nc1, d2 = ci[2] - nc2;
          double s = d0 * d0;
          s = s + d1 * d1;
          s = s + d2 * d2;
          if (s <= rc2) {
            if (cnt < cap) {
              jidx[cnt] = j;
              if (imgidx) imgidx[cnt] = img;
              ncoord[3 * cnt + 0] = nc0;
              ncoord[3 * cnt + 1] = nc1;
              ncoord[3 * cnt + 2] = nc2;
              dist2[cnt] = s;
            }
            ++cnt;
          }
        }
        ++img;
      }
  return cnt;
}

/* ------------------------------------------------------------------------- */
/* per-centre neighbour storage */

typedef struct {
  long n, cap;
  int* j;
  double* xyz; /* image coordinates (neighCoords) */
  double* d2;
} nblist;

static void nb_free(nblist* l) {
  free(l->j);
  free(l->xyz);
  free(l->d2);
  memset(l, 0, sizeof(*l));
}

static void nb_reserve(nblist* l, long cap) {
  if (cap <= l->cap) return;
  l->j = (int*)realloc(l->j, sizeof(int) * cap);
  l->xyz = (double*)realloc(l->xyz, sizeof(double) * 3 * cap);
  l->d2 = (double*)realloc(l->d2, sizeof(double) * cap);
  l->cap = cap;
}

static void nb_build(nblist* l, int n, const double* xyz, const double* lat, int periodic,
                     double rc, int i) {
  if (l->cap < 64) nb_reserve(l, 64);
  long c = orc_neighbours(n, xyz, lat, periodic, rc, i, l->cap, l->j, NULL, l->xyz, l->d2);
  if (c > l->cap) {
    nb_reserve(l, c);
    c = orc_neighbours(n, xyz, lat, periodic, rc, i, l->cap, l->j, NULL, l->xyz, l->d2);
  }
  l->n = c;
}

/* filtered list of one G-function (buildGFunctionNeighborlists, acsf.F90:943-1066) */
typedef struct {
  long n, cap;
  int* j;
  double* xyz;
  double* r;
  double* q; /* atom-ID prefactors of the neighbours */
} flist;

static void fl_free(flist* l) {
  free(l->j);
  free(l->xyz);
  free(l->r);
  free(l->q);
  memset(l, 0, sizeof(*l));
}

static void fl_filter(flist* out, const nblist* full, double rc, int z, int centre,
                      const int* atnum, const double* ext, int next, int atomid, int flags) {
  if (out->cap < full->n + 1) {
    long cap = full->n + 1;
    out->j = (int*)realloc(out->j, sizeof(int) * cap);
    out->xyz = (double*)realloc(out->xyz, sizeof(double) * 3 * cap);
    out->r = (double*)realloc(out->r, sizeof(double) * cap);
    out->q = (double*)realloc(out->q, sizeof(double) * cap);
    out->cap = cap;
  }
  const double rc2 = rc * rc;
  long c = 0;
  for (long e = 0; e < full->n; ++e) {
    int j = full->j[e];
    if (!(full->d2[e] <= rc2)) continue;
    if (z != 0 && atnum[j] != z) {
      if (!((flags & ORC_COMPAT_SELFIMG) && j == centre)) continue;
    }
    out->j[c] = j;
    out->xyz[3 * c + 0] = full->xyz[3 * e + 0];
    out->xyz[3 * c + 1] = full->xyz[3 * e + 1];
    out->xyz[3 * c + 2] = full->xyz[3 * e + 2];
    out->r[c] = sqrt(full->d2[e]); /* dynneighlist.F90:311 */
    out->q[c] = atomid > 0 ? ext[(long)j * next + (atomid - 1)] : 1.0;
    ++c;
  }
  out->n = c;
}

/* ------------------------------------------------------------------------- */
/* cutoff functions: acsf.F90:1181-1283 */

static inline double fc_wo(double rr, double q1, double q2, double rc) {
  return 0.5 * q1 * q2 * (cos(ORC_PI * rr / rc) + 1.0);
}

static inline double fc_chk(double rr, double q1, double q2, double rc) {
  if (rr > rc) return 0.0;
  return 0.5 * q1 * q2 * (cos(ORC_PI * rr / rc) + 1.0);
}

static inline double dfc_wo(double rr, double q1, double q2, double rc) {
  /* deriv = 1 */
  return 0.5 * q1 * q2 * (ORC_PI / rc) * cos((ORC_PI * rr / rc) + 0.5 * 1.0 * ORC_PI);
}

static inline double dot3(const double* a, const double* b) {
  double s = a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

static inline double norm3(const double* a) { return sqrt(dot3(a, a)); }

/* theta, acsf.F90:1159-1176 */
static inline double theta(const double* c, const double* n1, const double* n2, double r1,
                           double r2) {
  double a[3] = {n1[0] - c[0], n1[1] - c[1], n1[2] - c[2]};
  double b[3] = {n2[0] - c[0], n2[1] - c[1], n2[2] - c[2]};
  return acos(dot3(a, b) / (r1 * r2 + 1e-13));
}

/* ------------------------------------------------------------------------- */
/* G-function values: g1 :1287-1310, g2 :1314-1343, g3 :1347-1373, g4 :1377-1436, g5 :1440-1492 */

static double gvalue(const orc_gfuncs* gf, int a, const double* ci, double qi, const flist* l1,
                     const flist* l2, int flags) {
  const int ty = gf->type[a];
  const double rc = gf->rcut[a];
  double s = 0.0;
  if (ty == 1) {
    for (long e = 0; e < l1->n; ++e) s = s + fc_chk(l1->r[e], qi, l1->q[e], rc);
    return s;
  }
  if (ty == 2) {
    const double eta = gf->eta[a], rs = gf->rs[a];
    for (long e = 0; e < l1->n; ++e) {
      double dr = l1->r[e] - rs;
      s = s + exp(-eta * (dr * dr)) * fc_chk(l1->r[e], qi, l1->q[e], rc);
    }
    return s;
  }
  if (ty == 3) {
    const double kappa = gf->kappa[a];
    for (long e = 0; e < l1->n; ++e)
      s = s + cos(kappa * l1->r[e]) * fc_chk(l1->r[e], qi, l1->q[e], rc);
    return s;
  }
  const double xi = gf->xi[a], lam = gf->lam[a], eta = gf->eta[a];
  if (l1->n == 0 && l2->n == 0) return 0.0;
  for (long j = 0; j < l1->n; ++j) {
    const double* nj = l1->xyz + 3 * j;
    const double rj = l1->r[j];
    for (long k = 0; k < l2->n; ++k) {
      const double* nk = l2->xyz + 3 * k;
      const double rk = l2->r[k];
      double cth;
      if (flags & ORC_UNREG_COS) {
        const double u[3] = {nj[0] - ci[0], nj[1] - ci[1], nj[2] - ci[2]};
        const double v[3] = {nk[0] - ci[0], nk[1] - ci[1], nk[2] - ci[2]};
        cth = (l1 == l2 && j == k) ? 1.0 : dot3(u, v) * (1.0 / rj) * (1.0 / rk);
      } else {
        cth = cos(theta(ci, nj, nk, rj, rk));
      }
      double base = pow(1.0 + lam * cth, xi);
      if (ty == 4) {
        double djk[3] = {nj[0] - nk[0], nj[1] - nk[1], nj[2] - nk[2]};
        double distjk = norm3(djk);
        s = s + base * exp(-eta * (rj * rj + rk * rk + distjk * distjk)) *
                    fc_wo(rj, qi, l1->q[j], rc) * fc_wo(rk, qi, l2->q[k], rc) *
                    fc_chk(distjk, l1->q[j], l2->q[k], rc);
      } else {
        s = s + base * exp(-eta * (rj * rj + rk * rk)) * fc_wo(rj, qi, l1->q[j], rc) *
                    fc_wo(rk, qi, l2->q[k], rc);
      }
    }
  }
  return s * pow(2.0, 1.0 - xi);
}

/* ------------------------------------------------------------------------- */
/* G-function gradients w.r.t. the neighbour entries of list 1: gFuncGrad, acsf.F90:1496-1670.
 * grad[3*e] is dG/d(position of entry e of l1).  sameLists selects the Z1 == Z2 branch. */

static void tmatr(const double* u, double nrm, double T[3][3]) {
  /* T_matr, acsf.F90:1674-1697 */
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) T[a][b] = 0.0;
  for (int a = 0; a < 3; ++a) {
    T[a][a] = 1.0 - u[a] * u[a];
    for (int b = a + 1; b < 3; ++b) {
      T[a][b] = -u[a] * u[b];
      T[b][a] = T[a][b];
    }
  }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) T[a][b] = T[a][b] / nrm;
}

static void gradient(const orc_gfuncs* gf, int a, const double* ci, double qi, const flist* l1,
                     const flist* l2, int sameLists, double* grad) {
  const int ty = gf->type[a];
  const double rc = gf->rcut[a];
  for (long e = 0; e < 3 * l1->n; ++e) grad[e] = 0.0;
  if (ty == 1 || ty == 2 || ty == 3) {
    for (long e = 0; e < l1->n; ++e) {
      const double r = l1->r[e];
      double sc;
      if (ty == 1) {
        sc = dfc_wo(r, qi, l1->q[e], rc);
      } else if (ty == 2) {
        const double eta = gf->eta[a], rs = gf->rs[a];
        sc = (dfc_wo(r, qi, l1->q[e], rc) - fc_wo(r, qi, l1->q[e], rc) * 2.0 * eta * (r - rs)) *
             exp(-eta * ((r - rs) * (r - rs)));
      } else {
        const double kappa = gf->kappa[a];
        sc = cos(kappa * r) * dfc_wo(r, qi, l1->q[e], rc) -
             sin(kappa * r) * fc_wo(r, qi, l1->q[e], rc) * kappa;
      }
      for (int c = 0; c < 3; ++c) grad[3 * e + c] = sc * (l1->xyz[3 * e + c] - ci[c]) / r;
    }
    return;
  }
  const double eta = gf->eta[a], lam = gf->lam[a], xi = gf->xi[a];
  const long n2 = l2->n;
  double* u2 = (double*)malloc(sizeof(double) * 3 * (n2 + 1));
  double* fc2 = (double*)malloc(sizeof(double) * (n2 + 1));
  double* sq2 = (double*)malloc(sizeof(double) * (n2 + 1));
  for (long j = 0; j < n2; ++j) {
    for (int c = 0; c < 3; ++c) u2[3 * j + c] = (l2->xyz[3 * j + c] - ci[c]) / l2->r[j];
    sq2[j] = l2->r[j] * l2->r[j];
    fc2[j] = 0.0;
    if (l2->r[j] > rc) continue;
    fc2[j] = fc_wo(l2->r[j], qi, l2->q[j], rc);
  }
  if (sameLists) {
    for (long k = 0; k < l1->n; ++k) {
      double T[3][3];
      tmatr(u2 + 3 * k, l1->r[k], T);
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) T[p][q] = T[p][q] * lam * xi;
      const double dfc_ik =
          -2.0 * eta * l1->r[k] * fc2[k] + dfc_wo(l1->r[k], qi, l1->q[k], rc);
      double pp = 1.0;
      if (ty == 4) pp = l1->q[k] * l2->q[k];
      const double diag = pow(1.0 + lam, xi) * fc2[k] * dfc_ik * exp(-2.0 * eta * sq2[k]);
      for (int c = 0; c < 3; ++c) grad[3 * k + c] += diag * u2[3 * k + c] * pp;
      for (long j = 0; j < n2; ++j) {
        if (j == k) continue;
        const double aijk = 1.0 + lam * dot3(u2 + 3 * j, u2 + 3 * k);
        double tv[3];
        for (int p = 0; p < 3; ++p) {
          double s = 0.0;
          for (int q = 0; q < 3; ++q) s += T[p][q] * u2[3 * j + q];
          tv[p] = s;
        }
        if (ty == 5) {
          const double pre = fc2[j] * exp(-eta * (sq2[k] + sq2[j])) * pow(aijk, xi - 1.0);
          for (int c = 0; c < 3; ++c)
            grad[3 * k + c] += pre * (fc2[k] * tv[c] + aijk * dfc_ik * u2[3 * k + c]);
          continue;
        }
        double ujk[3] = {l1->xyz[3 * k + 0] - l2->xyz[3 * j + 0],
                         l1->xyz[3 * k + 1] - l2->xyz[3 * j + 1],
                         l1->xyz[3 * k + 2] - l2->xyz[3 * j + 2]};
        const double djk = norm3(ujk);
        if (djk > rc) continue;
        for (int c = 0; c < 3; ++c) ujk[c] = ujk[c] / djk;
        const double fcjk = fc_wo(djk, l1->q[k], l2->q[j], rc);
        const double dfcjk = -2.0 * eta * djk * fcjk + dfc_wo(djk, l1->q[k], l2->q[j], rc);
        const double pre =
            fc2[j] * exp(-eta * (sq2[k] + sq2[j] + djk * djk)) * pow(aijk, xi - 1.0);
        for (int c = 0; c < 3; ++c)
          grad[3 * k + c] += pre * (fc2[k] * fcjk * tv[c] +
                                    aijk * (fcjk * dfc_ik * u2[3 * k + c] + fc2[k] * dfcjk * ujk[c]));
      }
    }
    const double f = pow(2.0, 2.0 - xi);
    for (long e = 0; e < 3 * l1->n; ++e) grad[e] = grad[e] * f;
  } else {
    for (long k = 0; k < l1->n; ++k) {
      const double rk = l1->r[k];
      const double fcik = fc_wo(rk, qi, l1->q[k], rc);
      const double dfcik = -2.0 * eta * rk * fcik + dfc_wo(rk, qi, l1->q[k], rc);
      double uik[3];
      for (int c = 0; c < 3; ++c) uik[c] = (l1->xyz[3 * k + c] - ci[c]) / rk;
      double T[3][3];
      tmatr(uik, rk, T);
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) T[p][q] = T[p][q] * lam * xi;
      for (long j = 0; j < n2; ++j) {
        const double aijk = 1.0 + lam * dot3(u2 + 3 * j, uik);
        double tv[3];
        for (int p = 0; p < 3; ++p) {
          double s = 0.0;
          for (int q = 0; q < 3; ++q) s += T[p][q] * u2[3 * j + q];
          tv[p] = s;
        }
        if (ty == 5) {
          const double pre = fc2[j] * exp(-eta * (rk * rk + sq2[j])) * pow(aijk, xi - 1.0);
          for (int c = 0; c < 3; ++c)
            grad[3 * k + c] += pre * (fcik * tv[c] + aijk * dfcik * uik[c]);
          continue;
        }
        double ujk[3] = {l1->xyz[3 * k + 0] - l2->xyz[3 * j + 0],
                         l1->xyz[3 * k + 1] - l2->xyz[3 * j + 1],
                         l1->xyz[3 * k + 2] - l2->xyz[3 * j + 2]};
        const double djk = norm3(ujk);
        if (djk > rc) continue;
        for (int c = 0; c < 3; ++c) ujk[c] = ujk[c] / djk;
        const double fcjk = fc_wo(djk, l1->q[k], l2->q[j], rc);
        const double dfcjk = -2.0 * eta * djk * fcjk + dfc_wo(djk, l1->q[k], l2->q[j], rc);
        const double pre =
            fc2[j] * exp(-eta * (rk * rk + sq2[j] + djk * djk)) * pow(aijk, xi - 1.0);
        for (int c = 0; c < 3; ++c)
          grad[3 * k + c] += pre * (fcik * fcjk * tv[c] +
                                    aijk * (fcjk * dfcik * uik[c] + fcik * dfcjk * ujk[c]));
      }
    }
    const double f = pow(2.0, 1.0 - xi);
    for (long e = 0; e < 3 * l1->n; ++e) grad[e] = grad[e] * f;
  }
  free(u2);
  free(fc2);
  free(sq2);
}

/* ------------------------------------------------------------------------- */
/* Driver over one structure.  what: 1 = values, 2 = dense derivatives, 3 = adjoint forces. */

static double max_rcut(const orc_gfuncs* gf) {
  double m = 0.0;
  for (int a = 0; a < gf->nG; ++a)
    if (gf->rcut[a] > m) m = gf->rcut[a];
  return m;
}

static int run_structure(int what, int n, const double* xyz, const double* lat, int periodic,
                         const int* atnum, const double* ext, int next, const orc_gfuncs* gf,
                         int flags, double* G, double* dG, const double* cG, double* F) {
  nblist full;
  memset(&full, 0, sizeof(full));
  flist l1, l2;
  memset(&l1, 0, sizeof(l1));
  memset(&l2, 0, sizeof(l2));
  double* grad = NULL;
  long gradcap = 0;
  const double rcmax = max_rcut(gf);
  const int nG = gf->nG;
  for (int i = 0; i < n; ++i) {
    const double* ci = xyz + 3 * i;
    if (!(flags & ORC_REFSHAPE)) nb_build(&full, n, xyz, lat, periodic, rcmax, i);
    for (int a = 0; a < nG; ++a) {
      if (flags & ORC_REFSHAPE) nb_build(&full, n, xyz, lat, periodic, gf->rcut[a], i);
      const int resolved = !(gf->z1[a] == 0 && gf->z2[a] == 0);
      const int angular = gf->type[a] >= 4;
      const int atomid = gf->atomid[a];
      const double qi = atomid > 0 ? ext[(long)i * next + (atomid - 1)] : 1.0;
      fl_filter(&l1, &full, gf->rcut[a], resolved ? gf->z1[a] : 0, i, atnum, ext, next, atomid,
                flags);
      const flist* pl2 = &l1;
      if (resolved && angular) {
        fl_filter(&l2, &full, gf->rcut[a], gf->z2[a], i, atnum, ext, next, atomid, flags);
        pl2 = &l2;
      }
      if (what == 1) {
        G[(long)i * nG + a] = gvalue(gf, a, ci, qi, &l1, pl2, flags);
        continue;
      }
      const int same = (gf->z1[a] == gf->z2[a]);
      for (int pass = 0; pass < ((angular && !same) ? 2 : 1); ++pass) {
        const flist* la = pass == 0 ? &l1 : pl2;
        const flist* lb = pass == 0 ? pl2 : &l1;
        if (gradcap < 3 * la->n + 3) {
          gradcap = 3 * la->n + 3;
          grad = (double*)realloc(grad, sizeof(double) * gradcap);
        }
        gradient(gf, a, ci, qi, la, lb, same || !angular, grad);
        if (what == 2) {
          /* dense scatter: acsf.F90:918-933 */
          double* slot_i = dG + (((long)i * n + i) * nG + a) * 3;
          if (flags & ORC_COMPAT_SCATTER) {
            for (long e = 0; e < la->n; ++e) {
              double* s = dG + (((long)la->j[e] * n + i) * nG + a) * 3;
              for (int c = 0; c < 3; ++c) s[c] = grad[3 * e + c];
            }
            double sum[3] = {0.0, 0.0, 0.0};
            for (long e = 0; e < la->n; ++e) {
              double* s = dG + (((long)la->j[e] * n + i) * nG + a) * 3;
              for (int c = 0; c < 3; ++c) sum[c] += s[c];
            }
            if (pass == 0)
              for (int c = 0; c < 3; ++c) slot_i[c] = -sum[c];
            else
              for (int c = 0; c < 3; ++c) slot_i[c] = slot_i[c] - sum[c];
          } else {
            for (long e = 0; e < la->n; ++e) {
              double* s = dG + (((long)la->j[e] * n + i) * nG + a) * 3;
              for (int c = 0; c < 3; ++c) {
                s[c] += grad[3 * e + c];
                slot_i[c] -= grad[3 * e + c];
              }
            }
          }
        } else {
          /* adjoint: F_k -= sum_i sum_a cG[i,a] dG_ia/dr_k, image sums exact */
          const double cw = cG[(long)i * nG + a];
          if (cw == 0.0) continue;
          for (long e = 0; e < la->n; ++e) {
            const int j = la->j[e];
            for (int c = 0; c < 3; ++c) {
              F[3 * j + c] -= cw * grad[3 * e + c];
              F[3 * i + c] += cw * grad[3 * e + c];
            }
          }
        }
      }
    }
  }
  nb_free(&full);
  fl_free(&l1);
  fl_free(&l2);
  free(grad);
  return 0;
}

/* ACSF values of one structure: TAcsf_calculate -> iGeoAcsf, acsf.F90:540-639,797-866 */
int orc_acsf(int n, const double* xyz, const double* lat, int periodic, const int* atnum,
             const double* ext, int next, const orc_gfuncs* gf, int flags, double* G) {
  for (long q = 0; q < (long)n * gf->nG; ++q) G[q] = 0.0;
  return run_structure(1, n, xyz, lat, periodic, atnum, ext, next, gf, flags, G, NULL, NULL, NULL);
}

/* Dense ACSF derivatives: TAcsf_calculatePrime -> iGeoAcsfGrad, acsf.F90:643-717,870-939 */
int orc_acsf_prime(int n, const double* xyz, const double* lat, int periodic, const int* atnum,
                   const double* ext, int next, const orc_gfuncs* gf, int flags, double* dG) {
  for (long q = 0; q < (long)n * n * gf->nG * 3; ++q) dG[q] = 0.0;
  return run_structure(2, n, xyz, lat, periodic, atnum, ext, next, gf, flags, NULL, dG, NULL,
                       NULL);
}

/* Sparse adjoint of the same derivatives: F[k] -= sum_{i,a} cG[i,a] dG[k][i][a] (no dense array). */
int orc_acsf_adjoint(int n, const double* xyz, const double* lat, int periodic, const int* atnum,
                     const double* ext, int next, const orc_gfuncs* gf, int flags,
                     const double* cG, double* F) {
  for (long q = 0; q < 3L * n; ++q) F[q] = 0.0;
  return run_structure(3, n, xyz, lat, periodic, atnum, ext, next, gf, flags, NULL, NULL, cG, F);
}

/* ------------------------------------------------------------------------- */
/* Networks: network.F90, layer.F90, transfer.F90 */

/* activation ids */
enum { ACT_TANH = 0, ACT_SIGMOID, ACT_RELU, ACT_LRELU, ACT_SOFTPLUS, ACT_BENT, ACT_ATAN,
       ACT_GAUSSIAN, ACT_HEAVISIDE, ACT_LINEAR };

/* transfer.F90:54-342 */
static double act_f(int id, double x) {
  switch (id) {
    case ACT_GAUSSIAN: return exp(-(x * x));
    case ACT_SOFTPLUS: return log(1.0 + exp(x));
    case ACT_BENT: return (sqrt(x * x + 1.0) - 1.0) / 2.0 + x;
    case ACT_RELU: return x > 0.0 ? x : 0.0;
    case ACT_LRELU: return (0.01 * x > x) ? 0.01 * x : x;
    case ACT_SIGMOID: return 1.0 / (1.0 + exp(-x));
    case ACT_HEAVISIDE: return x > 0.0 ? 1.0 : 0.0;
    case ACT_ATAN: return atan(x);
    case ACT_LINEAR: return x;
    case ACT_TANH:
    default: return tanh(x);
  }
}

static double act_d(int id, double x) {
  switch (id) {
    case ACT_GAUSSIAN: return -2.0 * x * exp(-(x * x));
    case ACT_SOFTPLUS: return 1.0 / (1.0 + exp(-x));
    case ACT_BENT: return x / (2.0 * sqrt(x * x + 1.0)) + 1.0;
    case ACT_RELU: return x >= 0.0 ? 1.0 : 0.0;
    case ACT_LRELU: return x >= 0.0 ? 1.0 : 0.01;
    case ACT_SIGMOID: {
      double s = 1.0 / (1.0 + exp(-x));
      return s * (1.0 - s);
    }
    case ACT_HEAVISIDE: return 0.0;
    case ACT_ATAN: return 1.0 / (x * x + 1.0);
    case ACT_LINEAR: return 1.0;
    case ACT_TANH:
    default: {
      double t = tanh(x);
      return 1.0 - t * t;
    }
  }
}

/* offsets into the serialised parameter vector (network.F90:397-427, nestedtypes.F90:428-470) */
static long w_off(int nl, const int* dims, int l) {
  long o = 0;
  for (int q = 0; q < l; ++q) o += (long)dims[q] * dims[q + 1];
  (void)nl;
  return o;
}

long orc_nparams(int nl, const int* dims) {
  long nw = 0, nb = 0;
  for (int q = 0; q < nl - 1; ++q) nw += (long)dims[q] * dims[q + 1];
  nw += dims[nl - 1];
  for (int q = 0; q < nl; ++q) nb += dims[q];
  return nw + nb;
}

static long b_off(int nl, const int* dims, int l) {
  long o = 0;
  for (int q = 0; q < nl - 1; ++q) o += (long)dims[q] * dims[q + 1];
  o += dims[nl - 1];
  for (int q = 0; q < l; ++q) o += dims[q];
  return o;
}

/* fprop (network.F90:146-180) keeping pre-activations z and activations a per layer.
 * za: [sum(dims)] each for z and a, layer-major. */
static void net_fprop(int nl, const int* dims, int act, const double* p, const double* x,
                      double* z, double* a) {
  long off = 0;
  for (int q = 0; q < dims[0]; ++q) {
    z[q] = 0.0;
    a[q] = x[q];
  }
  for (int l = 1; l < nl; ++l) {
    const int din = dims[l - 1], dout = dims[l];
    const double* W = p + w_off(nl, dims, l - 1); /* ww(in, out) at W[out*din + in] */
    const double* b = p + b_off(nl, dims, l);
    const double* ain = a + off;
    double* zo = z + off + din;
    double* ao = a + off + din;
    const int id = (l == nl - 1) ? ACT_LINEAR : act;
    for (int o = 0; o < dout; ++o) {
      double s = 0.0;
      for (int in = 0; in < din; ++in) s += W[(long)o * din + in] * ain[in];
      zo[o] = s + b[o];
      ao[o] = act_f(id, zo[o]);
    }
    off += din;
  }
}

/* iPredict for many atoms of mixed species: bpnn.F90:867-900 */
int orc_bpnn_predict(int natom, int nl, const int* dims, const int* act, const double* params,
                     long nparams, const int* species, const double* feats, double* out) {
  long tot = 0;
  for (int l = 0; l < nl; ++l) tot += dims[l];
  double* z = (double*)malloc(sizeof(double) * tot);
  double* a = (double*)malloc(sizeof(double) * tot);
  const int nin = dims[0], nout = dims[nl - 1];
  for (int i = 0; i < natom; ++i) {
    const int sp = species[i];
    net_fprop(nl, dims, act[sp], params + (long)sp * nparams, feats + (long)i * nin, z, a);
    for (int o = 0; o < nout; ++o) out[(long)i * nout + o] = a[tot - nout + o];
  }
  free(z);
  free(a);
  return 0;
}

/* forward-mode Jacobian fdevi (network.F90:183-244): J[i][o][in] = d out_o / d x_in */
int orc_bpnn_jacobian(int natom, int nl, const int* dims, const int* act, const double* params,
                      long nparams, const int* species, const double* feats, double* J) {
  long tot = 0;
  int dmax = 0;
  for (int l = 0; l < nl; ++l) {
    tot += dims[l];
    if (dims[l] > dmax) dmax = dims[l];
  }
  const int nin = dims[0], nout = dims[nl - 1];
  double* z = (double*)malloc(sizeof(double) * tot);
  double* a = (double*)malloc(sizeof(double) * tot);
  double* cur = (double*)malloc(sizeof(double) * dmax * nin);
  double* nxt = (double*)malloc(sizeof(double) * dmax * nin);
  for (int i = 0; i < natom; ++i) {
    const int sp = species[i];
    const double* p = params + (long)sp * nparams;
    net_fprop(nl, dims, act[sp], p, feats + (long)i * nin, z, a);
    for (int r = 0; r < nin; ++r)
      for (int c = 0; c < nin; ++c) cur[r * nin + c] = (r == c) ? 1.0 : 0.0;
    long off = 0;
    for (int l = 1; l < nl; ++l) {
      const int din = dims[l - 1], dout = dims[l];
      const double* W = p + w_off(nl, dims, l - 1);
      const double* zo = z + off + din;
      const int id = (l == nl - 1) ? ACT_LINEAR : act[sp];
      for (int o = 0; o < dout; ++o) {
        const double dv = act_d(id, zo[o]);
        for (int c = 0; c < nin; ++c) {
          double s = 0.0;
          for (int in = 0; in < din; ++in) s += W[(long)o * din + in] * cur[in * nin + c];
          nxt[o * nin + c] = dv * s;
        }
      }
      memcpy(cur, nxt, sizeof(double) * dout * nin);
      off += din;
    }
    memcpy(J + (long)i * nout * nin, cur, sizeof(double) * nout * nin);
  }
  free(z);
  free(a);
  free(cur);
  free(nxt);
  return 0;
}

/* Per-structure training gradient: TBpnn_sysTrain (bpnn.F90:610-704) with bprop
 * (network.F90:248-296) and msGradients (loss.F90:251-264).  grad[nspecies][nparams] is
 * ACCUMULATED with factor dpweight (updateGradients, bpnn.F90:436-451).  Predictions of every
 * atom are written to pred[natom][nout]. */
int orc_sys_train(int natom, int nl, const int* dims, const int* act, const double* params,
                  long nparams, const int* species, const double* feats, int nglobal,
                  const double* gtargets, const double* atargets /*[natom][nout-nglobal]*/,
                  const double* atomicweights, double dpweight, double* pred, double* grad) {
  long tot = 0;
  int dmax = 0;
  for (int l = 0; l < nl; ++l) {
    tot += dims[l];
    if (dims[l] > dmax) dmax = dims[l];
  }
  const int nin = dims[0], nout = dims[nl - 1];
  const int natomic = nout - nglobal;
  double* z = (double*)malloc(sizeof(double) * tot * natom);
  double* a = (double*)malloc(sizeof(double) * tot * natom);
  double* lg = (double*)malloc(sizeof(double) * nout);
  double* dcur = (double*)malloc(sizeof(double) * dmax);
  double* dnxt = (double*)malloc(sizeof(double) * dmax);
  double* gsum = (double*)calloc(nglobal > 0 ? nglobal : 1, sizeof(double));
  for (int i = 0; i < natom; ++i) {
    net_fprop(nl, dims, act[species[i]], params + (long)species[i] * nparams, feats + (long)i * nin,
              z + tot * i, a + tot * i);
    for (int o = 0; o < nout; ++o) pred[(long)i * nout + o] = a[tot * i + tot - nout + o];
  }
  for (int o = 0; o < nglobal; ++o) {
    double s = 0.0;
    for (int i = 0; i < natom; ++i) s += pred[(long)i * nout + o];
    gsum[o] = 2.0 * (s - gtargets[o]);
  }
  for (int i = 0; i < natom; ++i) {
    const int sp = species[i];
    const double* p = params + (long)sp * nparams;
    double* g = grad + (long)sp * nparams;
    for (int o = 0; o < nglobal; ++o) lg[o] = gsum[o];
    for (int o = 0; o < natomic; ++o)
      lg[nglobal + o] = 2.0 * (pred[(long)i * nout + nglobal + o] - atargets[(long)i * natomic + o]);
    const double scale = atomicweights[i] / (double)natom * dpweight;
    const double* zi = z + tot * i;
    const double* ai = a + tot * i;
    /* layer offsets */
    long offs[64];
    offs[0] = 0;
    for (int l = 1; l < nl; ++l) offs[l] = offs[l - 1] + dims[l - 1];
    /* delta of the output layer (linear) */
    for (int o = 0; o < nout; ++o) dcur[o] = lg[o] * act_d(ACT_LINEAR, zi[offs[nl - 1] + o]);
    for (int l = nl - 1; l >= 1; --l) {
      const int din = dims[l - 1], dout = dims[l];
      /* db of layer l, dW of layer l-1 */
      double* gb = g + b_off(nl, dims, l);
      double* gw = g + w_off(nl, dims, l - 1);
      for (int o = 0; o < dout; ++o) gb[o] += dcur[o] * scale;
      for (int o = 0; o < dout; ++o)
        for (int in = 0; in < din; ++in) gw[(long)o * din + in] += ai[offs[l - 1] + in] * dcur[o] * scale;
      if (l == 1) break;
      const double* W = p + w_off(nl, dims, l - 1);
      for (int in = 0; in < din; ++in) {
        double s = 0.0;
        for (int o = 0; o < dout; ++o) s += W[(long)o * din + in] * dcur[o];
        dnxt[in] = s * act_d(act[sp], zi[offs[l - 1] + in]);
      }
      memcpy(dcur, dnxt, sizeof(double) * din);
    }
  }
  free(z);
  free(a);
  free(lg);
  free(dcur);
  free(dnxt);
  free(gsum);
  return 0;
}
