/*
 * CPU oracle (test infrastructure only): FP64 serial restatement of the
 * Genesis6 rigid-water interaction step.  Compiled with gcc -O2
 * -ffp-contract=off so that every + - * / sqrt is a single IEEE operation in
 * the same order as the reference's Fortran, which flang -O2 compiles for
 * baseline x86-64 without FMA.  See g6_oracle.h for the per-function
 * reference citations.  Not part of the product: the product library never
 * links or calls it.
 */
#include "g6_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- physical constants (PhysConst.F90:5-62) --------------------------- */
static const double Gear5_0 = 3.0 / 16.0, Gear5_1 = 251.0 / 360.0,
                    Gear5_3 = 11.0 / 18.0, Gear5_4 = 1.0 / 6.0,
                    Gear5_5 = 1.0 / 60.0;
static const double Gear4_0 = 251.0 / 720.0, Gear4_2 = 11.0 / 12.0,
                    Gear4_3 = 1.0 / 3.0, Gear4_4 = 1.0 / 24.0;
#define PI_T 3.1415926535897932
#define NA_ 6.0225e23
#define KB_ 1.380066e-23
#define CA_ 4.184
static double rgas(void) { return NA_ * KB_; }
static double coeff(void) { return CA_ * 332.17752; }
static const double I2J = 10.0;
static double J2I(void) { return 1.0 / 10.0; }
static double KX2X(void) { return 1.0 / 1e-3; }

/* x**k for small integer k as repeated multiplication */
static double ipow(double x, int k) {
  double r = x;
  for (int i = 1; i < k; i++) r = r * x;
  return r;
}

/* ---- models ------------------------------------------------------------ */
int g6o_model_get(int model, g6o_model *m) {
  memset(m, 0, sizeof(*m));
  if (model == G6O_TIP4P) {
    /* Rigid_TIP4P_Constructor, TIP4PModule.F90:70-97 (moly(5) unset = 0) */
    m->nsite = 5;
    m->Ixx = 1.75618411509879202e+00;
    m->Iyy = 6.10236519209664374e-01;
    m->Izz = 1.14594759588912742e+00;
    m->mass = 18.0;
    m->by[2] = 7.56950327263661182e-01;
    m->by[3] = -7.56950327263661182e-01;
    m->bz[0] = -6.50980307353661025e-02;
    m->bz[1] = 8.49019692646338919e-02;
    m->bz[2] = 5.20784245882928820e-01;
    m->bz[3] = 5.20784245882928820e-01;
    /* tip4p_setinteraction, TIP4PModule.F90:43-68 */
    m->q[1] = -1.04;
    m->q[2] = m->q[3] = +0.52;
    m->eps[0] = 0.15504166666666666666666 * CA_;
    m->sig[0] = 3.15357794197649532464;
    return m->nsite;
  }
  if (model == G6O_TIP5P) {
    /* Rigid_TIP5P_Constructor, RigidModule.F90:429-477 */
    const double theta = 104.52 * PI_T / 180.0, phi = 109.47 * PI_T / 180.0;
    const double l1 = 0.9572, l2 = 0.70, hmass = 1.0, omass = 16.0;
    m->nsite = 6;
    m->by[1] = l1 * sin(0.5 * theta);
    m->bz[1] = l1 * cos(0.5 * theta);
    m->by[2] = -l1 * sin(0.5 * theta);
    m->bz[2] = l1 * cos(0.5 * theta);
    m->bx[3] = l2 * sin(0.5 * phi);
    m->bz[3] = -l2 * cos(0.5 * phi);
    m->bx[4] = -l2 * sin(0.5 * phi);
    m->bz[4] = -l2 * cos(0.5 * phi);
    double zshift = m->bz[1] * 2.0 * hmass / (2.0 * hmass + omass);
    for (int i = 0; i < 5; i++) m->bz[i] = m->bz[i] - zshift;
    m->Ixx = omass * (m->by[0] * m->by[0] + m->bz[0] * m->bz[0]) +
             hmass * (m->by[1] * m->by[1] + m->bz[1] * m->bz[1]) * 2.0;
    m->Iyy = omass * (m->bz[0] * m->bz[0] + m->bx[0] * m->bx[0]) +
             hmass * (m->bz[1] * m->bz[1] + m->bx[1] * m->bx[1]) * 2.0;
    m->Izz = omass * (m->bx[0] * m->bx[0] + m->by[0] * m->by[0]) +
             hmass * (m->bx[1] * m->bx[1] + m->by[1] * m->by[1]) * 2.0;
    m->mass = 2.0 * hmass + omass;
    /* tip5p_setinteraction, TIP5PModule.F90:57-83 */
    m->q[1] = m->q[2] = 0.2410;
    m->q[3] = m->q[4] = -0.2410;
    m->eps[0] = 0.6694;
    m->sig[0] = 3.12;
    return m->nsite;
  }
  if (model == G6O_ST2) {
    /* Rigid_ST2_Constructor, ST2Module.F90:60-100 */
    const double theta = 109.47 * PI_T / 180.0, l1 = 1.0, l2 = 0.8;
    const double hmass = 1.0, omass = 16.0;
    m->nsite = 6;
    m->by[1] = l1 * sin(0.5 * theta);
    m->bz[1] = l1 * cos(0.5 * theta);
    m->by[2] = -l1 * sin(0.5 * theta);
    m->bz[2] = l1 * cos(0.5 * theta);
    m->bx[3] = l2 * sin(0.5 * theta);
    m->bz[3] = -l2 * cos(0.5 * theta);
    m->bx[4] = -l2 * sin(0.5 * theta);
    m->bz[4] = -l2 * cos(0.5 * theta);
    double zshift = m->bz[1] * 2.0 * hmass / (2.0 * hmass + omass);
    for (int i = 0; i < 5; i++) m->bz[i] = m->bz[i] - zshift;
    m->Ixx = omass * (m->by[0] * m->by[0] + m->bz[0] * m->bz[0]) +
             hmass * (m->by[1] * m->by[1] + m->bz[1] * m->bz[1]) * 2.0;
    m->Iyy = omass * (m->bz[0] * m->bz[0] + m->bx[0] * m->bx[0]) +
             hmass * (m->bz[1] * m->bz[1] + m->bx[1] * m->bx[1]) * 2.0;
    m->Izz = omass * (m->bx[0] * m->bx[0] + m->by[0] * m->by[0]) +
             hmass * (m->bx[1] * m->bx[1] + m->by[1] * m->by[1]) * 2.0;
    m->mass = 2.0 * hmass + omass;
    /* st2_setinteraction, ST2Module.F90:32-58.  The charges and eps are
     * default-REAL (single precision) literals in the reference
     * (ST2Module.F90:15-17), widened to REAL(8).                           */
    m->q[1] = m->q[2] = (double)0.2357f;
    m->q[3] = m->q[4] = (double)(-0.2357f);
    m->eps[0] = (double)0.52605e-21f * NA_ * 1e-3;
    m->sig[0] = 3.100;
    return m->nsite;
  }
  if (model == G6O_SPCE) {
    /* Rigid_SPCE_Constructor, SPCEModule.F90:48-80: O, H, H, CoM */
    const double theta = 109.5 * PI_T / 180.0, l1 = 1.0, hmass = 1.0, omass = 16.0;
    m->nsite = 4;
    m->by[1] = l1 * sin(0.5 * theta);
    m->bz[1] = l1 * cos(0.5 * theta);
    m->by[2] = -l1 * sin(0.5 * theta);
    m->bz[2] = l1 * cos(0.5 * theta);
    double zshift = m->bz[1] * 2.0 * hmass / (2.0 * hmass + omass);
    for (int i = 0; i < 3; i++) m->bz[i] = m->bz[i] - zshift;
    m->Ixx = omass * (m->by[0] * m->by[0] + m->bz[0] * m->bz[0]) +
             hmass * (m->by[1] * m->by[1] + m->bz[1] * m->bz[1]) * 2.0;
    m->Iyy = omass * (m->bz[0] * m->bz[0] + m->bx[0] * m->bx[0]) +
             hmass * (m->bz[1] * m->bz[1] + m->bx[1] * m->bx[1]) * 2.0;
    m->Izz = omass * (m->bx[0] * m->bx[0] + m->by[0] * m->by[0]) +
             hmass * (m->bx[1] * m->bx[1] + m->by[1] * m->by[1]) * 2.0;
    m->mass = 2.0 * hmass + omass;
    /* spce_setinteraction, SPCEModule.F90:25-46: O carries the charge and
     * the LJ site; eps in K converted by J2K (PhysConst.F90:56) */
    m->q[0] = -0.8476;
    m->q[1] = m->q[2] = +0.4238;
    m->eps[0] = 78.208 / 0.120273 * 1e-3;
    m->sig[0] = 3.166;
    return m->nsite;
  }
  if (model == G6O_NVDE) {
    /* Rigid_NvdE_Constructor, NvdEModule.F90:90-148: O, H, H, M, L, L, CoM.
     * The constructor shifts x by comz/mass (NvdEModule.F90:125), kept. */
    const double oh = 0.980, om = 0.230, ol = 0.8892;
    const double hoh = 108.00 * PI_T / 180.0, lol = 111.0 * PI_T / 180.0;
    const double ms[7] = {16.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0};
    double x[7] = {0, 0, 0, 0, 0, 0, 0}, y[7] = {0, 0, 0, 0, 0, 0, 0}, z[7] = {0, 0, 0, 0, 0, 0, 0};
    y[1] = oh * sin(0.5 * hoh);
    z[1] = oh * cos(0.5 * hoh);
    y[2] = -oh * sin(0.5 * hoh);
    z[2] = oh * cos(0.5 * hoh);
    z[3] = om;
    x[4] = ol * sin(0.5 * lol);
    z[4] = -ol * cos(0.5 * lol);
    x[5] = -ol * sin(0.5 * lol);
    z[5] = -ol * cos(0.5 * lol);
    double cx = 0, cy = 0, cz = 0, mass = 0;
    for (int i = 0; i < 6; i++) {
      cx = cx + x[i] * ms[i];
      cy = cy + y[i] * ms[i];
      cz = cz + z[i] * ms[i];
      mass = mass + ms[i];
    }
    const double mi = 1.0 / mass;
    cx = cz * mi;
    cy = cy * mi;
    cz = cz * mi;
    for (int i = 0; i < 6; i++) {
      x[i] = x[i] - cx;
      y[i] = y[i] - cy;
      z[i] = z[i] - cz;
    }
    m->nsite = 7;
    for (int i = 0; i < 7; i++) {
      m->bx[i] = x[i];
      m->by[i] = y[i];
      m->bz[i] = z[i];
      m->Ixx = m->Ixx + ms[i] * (y[i] * y[i] + z[i] * z[i]);
      m->Iyy = m->Iyy + ms[i] * (z[i] * z[i] + x[i] * x[i]);
      m->Izz = m->Izz + ms[i] * (x[i] * x[i] + y[i] * y[i]);
    }
    m->mass = mass;
    /* nvde_setinteraction, NvdEModule.F90:58-88 (parameters :10-21) */
    m->q[1] = m->q[2] = 0.477;
    m->q[3] = -0.866;
    m->q[4] = m->q[5] = -0.044;
    m->eps[0] = 85.9766 * rgas() * 1e-3;
    m->eps[1] = m->eps[2] = 13.8817 * rgas() * 1e-3;
    m->sig[0] = 3.115;
    m->sig[1] = m->sig[2] = 0.673;
    return m->nsite;
  }
  if (model == G6O_OPLSMEOH) {
    /* Rigid_OPLSMeOH_Constructor, RigidModule.F90:391-426: Me, O, H, CoM in
     * the plane z = 0, as given */
    const double ms[4] = {15.0, 16.0, 1.0, 0.0};
    const double x[4] = {0.0133, -0.0634, 0.8152, 0.0};
    const double y[4] = {0.7666, -0.6559, -1.0042, 0.0};
    m->nsite = 4;
    for (int i = 0; i < 4; i++) {
      m->bx[i] = x[i];
      m->by[i] = y[i];
      m->bz[i] = 0.0;
      m->Ixx = m->Ixx + ms[i] * (m->by[i] * m->by[i] + m->bz[i] * m->bz[i]);
      m->Iyy = m->Iyy + ms[i] * (m->bz[i] * m->bz[i] + m->bx[i] * m->bx[i]);
      m->Izz = m->Izz + ms[i] * (m->bx[i] * m->bx[i] + m->by[i] * m->by[i]);
      m->mass = m->mass + ms[i];
    }
    /* oplsmeoh_setinteraction, OPLSMeOHModule.F90:82-101 (parameters :41-49,
     * eps in kcal/mol times CA, PhysConst.F90:29-31) */
    m->q[0] = 0.265;
    m->q[1] = -0.700;
    m->q[2] = 0.435;
    m->eps[0] = 0.207 * CA_;
    m->eps[1] = 0.170 * CA_;
    m->sig[0] = 3.775;
    m->sig[1] = 3.070;
    return m->nsite;
  }
  return 0;
}

/* [WWMM, WWMH, WWHH, AA, BB, QMM, QHM, QHH, AAst2, BBst2, RL, RU] */
void g6o_constants(double *o) {
  const double C = coeff();
  o[0] = -1.04 * -1.04 * C * 100.0;   /* TIP4PConst.F90:11 */
  o[1] = -1.04 * 0.52 * C * 100.0;    /* TIP4PConst.F90:12 */
  o[2] = 0.52 * 0.52 * C * 100.0;     /* TIP4PConst.F90:13 */
  o[3] = 600000.0 * CA_ * 100.0;      /* TIP4PConst.F90:4 */
  o[4] = -610.0 * CA_ * 100.0;        /* TIP4PConst.F90:5 */
  g6o_model st2;
  g6o_model_get(G6O_ST2, &st2);
  const double qm = st2.q[3], qh = st2.q[1], eps = st2.eps[0], sig = st2.sig[0];
  o[5] = qm * qm * C * 100.0;         /* ST2Module.F90:121-123 */
  o[6] = qh * qm * C * 100.0;
  o[7] = qh * qh * C * 100.0;
  o[8] = 4.0 * eps * ipow(sig, 12) * 100.0;
  o[9] = -4.0 * eps * ipow(sig, 6) * 100.0;
  o[10] = 2.0160;                     /* ST2Module.F90:21-22 */
  o[11] = 3.1287;
}

/* ---- grid (GridModule.F90) --------------------------------------------- */
int g6o_grid_ndiv(double L, double req) { return (int)(L / req); }

int g6o_stencil_homo(const int nd[3], int *sx, int *sy, int *sz, int *n0p,
                     int *sx0, int *sy0, int *sz0) {
  int nmax[3], nmin[3];
  for (int a = 0; a < 3; a++) {
    nmax[a] = 1;
    nmin[a] = -1;
    if (nmax[a] > (nd[a] - 1) / 2) nmax[a] = (nd[a] - 1) / 2;
    if (nmin[a] < -(nd[a] / 2)) nmin[a] = -(nd[a] / 2);
  }
  int *mark = calloc((size_t)nd[0] * nd[1] * nd[2], sizeof(int));
#define MK(x, y, z) mark[((z) * nd[1] + (y)) * nd[0] + (x)]
  int n = 0, n0 = 0;
  for (int jz = nmin[2]; jz <= nmax[2]; jz++) {
    int pz = jz < 0 ? jz + nd[2] : jz, mz = pz == 0 ? 0 : nd[2] - pz;
    for (int jy = nmin[1]; jy <= nmax[1]; jy++) {
      int py = jy < 0 ? jy + nd[1] : jy, my = py == 0 ? 0 : nd[1] - py;
      for (int jx = nmin[0]; jx <= nmax[0]; jx++) {
        int px = jx < 0 ? jx + nd[0] : jx, mx = px == 0 ? 0 : nd[0] - px;
        if (MK(px, py, pz) == 0) {
          MK(mx, my, mz) += 1;
          MK(px, py, pz) += 1;
          if (MK(px, py, pz) != 1) {
            sx0[n0] = px; sy0[n0] = py; sz0[n0] = pz; n0++;
          } else {
            sx[n] = px; sy[n] = py; sz[n] = pz; n++;
          }
        }
      }
    }
  }
#undef MK
  free(mark);
  *n0p = n0;
  return n;
}

void g6o_address_assign(int n, const double *com, double L, const int nd[3],
                        int *cell) {
  const double invL = 1.0 / L;
  double ci[3];
  for (int a = 0; a < 3; a++) ci[a] = 1.0 / (L / nd[a]);
  for (int i = 0; i < n; i++) {
    int j[3];
    for (int a = 0; a < 3; a++) {
      double d = com[3 * i + a];
      double c = round(d * invL) * L;       /* box_renormalize */
      d = d - c;
      j[a] = (int)round(d * ci[a]);
      if (j[a] < 0) j[a] += nd[a];
    }
    cell[i] = (j[2] * nd[1] + j[1]) * nd[0] + j[0];
  }
}

/* minimum-image FP64 test of one candidate (InteractionModule.F90:454-474) */
static int compress_one(const double *com, int i, int j, double L, double invL,
                        double rr, double *o) {
  double d[3];
  for (int a = 0; a < 3; a++) {
    d[a] = com[3 * i + a] - com[3 * j + a];
    o[a] = round(d[a] * invL) * L;
    d[a] = d[a] - o[a];
  }
  double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  return dd < rr;
}

long g6o_pairlist(int n, const double *com, double L, double rc, double req,
                  long cap, long *npair0, int *pi, int *pj, double *ox,
                  double *oy, double *oz) {
  int nd[3];
  nd[0] = nd[1] = nd[2] = g6o_grid_ndiv(L, req);
  int sx[27], sy[27], sz[27], sx0[27], sy0[27], sz0[27], n0;
  int ns = g6o_stencil_homo(nd, sx, sy, sz, &n0, sx0, sy0, sz0);
  const int ng = nd[0] * nd[1] * nd[2];
  int *cell = malloc(sizeof(int) * n);
  g6o_address_assign(n, com, L, nd, cell);
  /* residents in molecule order (GridModule.F90:538-548) */
  int *cnt = calloc(ng + 1, sizeof(int)), *start = calloc(ng + 1, sizeof(int));
  int *res = malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) cnt[cell[i]]++;
  for (int c = 0; c < ng; c++) start[c + 1] = start[c] + cnt[c];
  memset(cnt, 0, sizeof(int) * (ng + 1));
  for (int i = 0; i < n; i++) res[start[cell[i]] + cnt[cell[i]]++] = i;
  const double invL = 1.0 / L, rr = rc * rc;
  long k0 = 0, k = 0;
  for (int part = 0; part < 2; part++) {
    const int nst = part == 0 ? ns : n0;
    const int *tx = part == 0 ? sx : sx0, *ty = part == 0 ? sy : sy0,
              *tz = part == 0 ? sz : sz0;
    for (int c1 = 0; c1 < ng; c1++) {
      const int ix = c1 % nd[0], iy = (c1 / nd[0]) % nd[1], iz = c1 / (nd[0] * nd[1]);
      for (int ii = start[c1]; ii < start[c1 + 1]; ii++) {
        const int i = res[ii];
        for (int s = 0; s < nst; s++) {
          const int c2 = (((iz + tz[s]) % nd[2]) * nd[1] + (iy + ty[s]) % nd[1]) * nd[0] +
                         (ix + tx[s]) % nd[0];
          for (int jj = start[c2]; jj < start[c2 + 1]; jj++) {
            const int j = res[jj];
            if (part == 1 && !(j > i)) continue;
            k0++;
            double o[3];
            if (compress_one(com, i, j, L, invL, rr, o)) {
              if (k >= cap) { k = -1; goto done; }
              pi[k] = i; pj[k] = j; ox[k] = o[0]; oy[k] = o[1]; oz[k] = o[2];
              k++;
            }
          }
        }
      }
    }
  }
done:
  *npair0 = k0;
  free(cell); free(cnt); free(start); free(res);
  return k;
}

long g6o_pairlist_alltoall(int n, const double *com, double L, double rc,
                           long cap, int *pi, int *pj, double *ox, double *oy,
                           double *oz) {
  const double invL = 1.0 / L, rr = rc * rc;
  long k = 0;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      double o[3];
      if (compress_one(com, i, j, L, invL, rr, o)) {
        if (k >= cap) return -1;
        pi[k] = i; pj[k] = j; ox[k] = o[0]; oy[k] = o[1]; oz[k] = o[2];
        k++;
      }
    }
  return k;
}

/* ---- cutoff (CutoffModule.F90:25-59, 220-278) --------------------------- */
void g6o_cutoff_func(double rin, double rout, double dd, double *hh, double *dh) {
  const double c1 = 1.0 / ipow(rin - rout, 5), c2 = c1 * 30.0;
  *hh = 0.0;
  *dh = 0.0;
  if (dd < rout * rout) {
    if (dd > rin * rin) {
      double dr = sqrt(dd), u = dr - rout, l = dr - rin, uu = u * u, ll = l * l;
      *hh = c1 * uu * u * (10.0 * ll - 5.0 * u * l + uu);
      *dh = -(c2 * uu * ll / dr);
    } else {
      *hh = 1.0;
      *dh = 0.0;
    }
  }
}

void g6o_cutoff_dimmer(long np, const int *pi, const int *pj, const double *com,
                       const double *ox, const double *oy, const double *oz,
                       double rin, double rout, double *er, double *fr) {
  for (long k = 0; k < np; k++) {
    const int i = pi[k], j = pj[k];
    double dx = com[3 * i] - com[3 * j], dy = com[3 * i + 1] - com[3 * j + 1],
           dz = com[3 * i + 2] - com[3 * j + 2];
    dx = dx - ox[k]; dy = dy - oy[k]; dz = dz - oz[k];
    g6o_cutoff_func(rin, rout, dx * dx + dy * dy + dz * dz, &er[k], &fr[k]);
  }
}

/* ---- TIP4P custom kernel (TIP4PModule.F90:187-785) ---------------------- */
void g6o_force_tip4p(long np, const int *pi, const int *pj, const double *ox,
                     const double *oy, const double *oz, const double *er,
                     const double *fr, const double *x, const double *y,
                     const double *z, double *fx, double *fy, double *fz,
                     double *ep1, double *vr) {
  double K[12];
  g6o_constants(K);
  const double WWMM = K[0], WWMH = K[1], WWHH = K[2], AA = K[3], BB = K[4];
  /* the nine Coulomb pairs in the order of the live (non-TEST12) code */
  static const int ci[9] = {1, 1, 1, 2, 3, 2, 2, 3, 3};
  static const int cj[9] = {1, 2, 3, 1, 1, 2, 3, 2, 3};
  const double cq[9] = {WWMM, WWMH, WWMH, WWMH, WWMH, WWHH, WWHH, WWHH, WWHH};
  for (int a = 0; a < 9; a++) vr[a] = 0.0;
  for (long k = 0; k < np; k++) {
    const int ii = pi[k] * 5, jj = pj[k] * 5;
    double dx00 = x[ii + 4], dy00 = y[ii + 4], dz00 = z[ii + 4];
    dx00 = dx00 - x[jj + 4]; dy00 = dy00 - y[jj + 4]; dz00 = dz00 - z[jj + 4];
    dx00 = dx00 - ox[k]; dy00 = dy00 - oy[k]; dz00 = dz00 - oz[k];
    const double hh = er[k], dh = fr[k];
    double dhx = dh * dx00, dhy = dh * dy00, dhz = dh * dz00;
    const double cx = -ox[k], cy = -oy[k], cz = -oz[k];
    double fi[5][3] = {{0}}, fj[5][3] = {{0}};
    double ep = 0.0;
    for (int t = 0; t < 9; t++) {
      const int a = ci[t], b = cj[t];
      double dx = x[ii + a] - x[jj + b] + cx, dy = y[ii + a] - y[jj + b] + cy,
             dz = z[ii + a] - z[jj + b] + cz;
      double dd = 1.0 / (dx * dx + dy * dy + dz * dz);
      double ri = sqrt(dd);
      double e0 = cq[t] * ri, f0 = cq[t] * ri * dd;
      ep = ep + e0;
      double gx = dx * f0, gy = dy * f0, gz = dz * f0;
      fi[a][0] += gx; fi[a][1] += gy; fi[a][2] += gz;
      fj[b][0] += gx; fj[b][1] += gy; fj[b][2] += gz;
    }
    {
      double dx = x[ii] - x[jj] + cx, dy = y[ii] - y[jj] + cy, dz = z[ii] - z[jj] + cz;
      double dd = 1.0 / (dx * dx + dy * dy + dz * dz);
      double d6 = dd * dd * dd, d12 = d6 * d6;
      double e0 = AA * d12 + BB * d6;
      double f0 = dd * (AA * 12.0 * d12 + BB * 6.0 * d6);
      ep = ep + e0;
      double gx = dx * f0, gy = dy * f0, gz = dz * f0;
      fi[0][0] += gx; fi[0][1] += gy; fi[0][2] += gz;
      fj[0][0] += gx; fj[0][1] += gy; fj[0][2] += gz;
    }
    for (int s = 0; s < 5; s++)
      for (int a = 0; a < 3; a++) { fi[s][a] = fi[s][a] * hh; fj[s][a] = fj[s][a] * hh; }
    dhx = dhx * ep; dhy = dhy * ep; dhz = dhz * ep;
    fi[4][0] += dhx; fi[4][1] += dhy; fi[4][2] += dhz;
    fj[4][0] += dhx; fj[4][1] += dhy; fj[4][2] += dhz;
    ep1[k] = ep * hh;
    double v[3];
    for (int a = 0; a < 3; a++)
      v[a] = fi[0][a] + fi[1][a] + fi[2][a] + fi[3][a] + fi[4][a];
    const double d00[3] = {dx00, dy00, dz00};
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) vr[3 * a + b] = vr[3 * a + b] + d00[a] * v[b];
    for (int s = 0; s < 5; s++) {
      fx[ii + s] += fi[s][0]; fy[ii + s] += fi[s][1]; fz[ii + s] += fi[s][2];
      fx[jj + s] -= fj[s][0]; fy[jj + s] -= fj[s][1]; fz[jj + s] -= fj[s][2];
    }
  }
}

/* ---- generic StdForce (StdInteraction.F90:63-630) ----------------------- */
typedef struct { int a, b, kind; double qq, aa, bb; } stdterm;
/* site-pair terms in the isite-outer, jsite-inner order with the reference's
 * coefficient arithmetic (StdInteraction.F90:201-221); kind 1=LJ+C, 2=C, 3=LJ */
static int std_terms(const g6o_model *m, stdterm *t) {
  int nt = 0;
  const double scale = 1.0;
  for (int a = 0; a < m->nsite; a++)
    for (int b = 0; b < m->nsite; b++) {
      double qq = m->q[a] * m->q[b] * coeff() * KX2X() * J2I() * scale;
      double aa = m->eps[a] * m->eps[b];
      stdterm s = {a, b, 0, qq, 0.0, 0.0};
      if (qq != 0.0 || aa != 0.0) {
        if (aa != 0.0) {
          double ljeps = sqrt(aa);
          double ljsig = (m->sig[a] + m->sig[b]) * 0.5;
          ljsig = ipow(ljsig, 6);
          s.aa = 4.0 * ljeps * ljsig * ljsig * KX2X() * J2I() * scale;
          s.bb = -4.0 * ljeps * ljsig * KX2X() * J2I() * scale;
        }
        s.kind = qq != 0.0 ? (aa != 0.0 ? 1 : 2) : 3;
        t[nt++] = s;
      }
    }
  return nt;
}

void g6o_force_std(int model, long np, const int *pi, const int *pj,
                   const double *ox, const double *oy, const double *oz,
                   const double *er, const double *fr, const double *x,
                   const double *y, const double *z, double *fx, double *fy,
                   double *fz, double *ep, double *vr) {
  g6o_model m;
  g6o_model_get(model, &m);
  const int ns = m.nsite;
  stdterm t[G6O_MAXSITE * G6O_MAXSITE];
  const int nt = std_terms(&m, t);
  double *fi = calloc((size_t)np * ns * 3 + 1, sizeof(double));
  double *fj = calloc((size_t)np * ns * 3 + 1, sizeof(double));
#define FI(k, s, a) fi[((size_t)(s) * np + (k)) * 3 + (a)]
#define FJ(k, s, a) fj[((size_t)(s) * np + (k)) * 3 + (a)]
  for (long k = 0; k < np; k++) ep[k] = 0.0;
  for (int u = 0; u < nt; u++) {
    const stdterm s = t[u];
    for (long k = 0; k < np; k++) {
      const int ii = pi[k] * ns + s.a, jj = pj[k] * ns + s.b;
      double dx = x[ii] - x[jj], dy = y[ii] - y[jj], dz = z[ii] - z[jj];
      dx = dx - ox[k]; dy = dy - oy[k]; dz = dz - oz[k];
      double dd = 1.0 / (dx * dx + dy * dy + dz * dz);
      double e0, f0;
      if (s.kind == 1) {
        double ri = sqrt(dd);
        e0 = s.qq * ri;
        f0 = s.qq * ri * dd;
        double d6 = dd * dd * dd, d12 = d6 * d6;
        e0 = e0 + s.aa * d12 + s.bb * d6;
        f0 = f0 + dd * (s.aa * 12.0 * d12 + s.bb * 6.0 * d6);
      } else if (s.kind == 2) {
        double ri = sqrt(dd);
        e0 = s.qq * ri;
        f0 = s.qq * ri * dd;
      } else {
        double d6 = dd * dd * dd, d12 = d6 * d6;
        e0 = s.aa * d12 + s.bb * d6;
        f0 = dd * (s.aa * 12.0 * d12 + s.bb * 6.0 * d6);
      }
      double gx = dx * f0, gy = dy * f0, gz = dz * f0;
      FI(k, s.a, 0) += gx; FI(k, s.a, 1) += gy; FI(k, s.a, 2) += gz;
      FJ(k, s.b, 0) += gx; FJ(k, s.b, 1) += gy; FJ(k, s.b, 2) += gz;
      ep[k] = ep[k] + e0;
    }
  }
  /* smoothing (491-514), CoM term (515-536) */
  for (int l = 0; l < ns; l++)
    for (long k = 0; k < np; k++)
      for (int a = 0; a < 3; a++) FI(k, l, a) = FI(k, l, a) * er[k];
  for (int l = 0; l < ns; l++)
    for (long k = 0; k < np; k++)
      for (int a = 0; a < 3; a++) FJ(k, l, a) = FJ(k, l, a) * er[k];
  const int c = ns - 1;
  for (long k = 0; k < np; k++) {
    const int ii = pi[k] * ns + c, jj = pj[k] * ns + c;
    double d0 = x[ii] - x[jj] - ox[k], d1 = y[ii] - y[jj] - oy[k],
           d2 = z[ii] - z[jj] - oz[k];
    double dhx = fr[k] * d0 * ep[k], dhy = fr[k] * d1 * ep[k], dhz = fr[k] * d2 * ep[k];
    FI(k, c, 0) += dhx; FI(k, c, 1) += dhy; FI(k, c, 2) += dhz;
    FJ(k, c, 0) += dhx; FJ(k, c, 1) += dhy; FJ(k, c, 2) += dhz;
  }
  /* scatter, TEST22 order (537-560) */
  for (int l = 0; l < ns; l++)
    for (long k = 0; k < np; k++) {
      const int ii = pi[k] * ns + l;
      fx[ii] = fx[ii] + FI(k, l, 0); fy[ii] = fy[ii] + FI(k, l, 1); fz[ii] = fz[ii] + FI(k, l, 2);
    }
  for (int l = 0; l < ns; l++)
    for (long k = 0; k < np; k++) {
      const int jj = pj[k] * ns + l;
      fx[jj] = fx[jj] - FJ(k, l, 0); fy[jj] = fy[jj] - FJ(k, l, 1); fz[jj] = fz[jj] - FJ(k, l, 2);
    }
  for (long k = 0; k < np; k++) ep[k] = ep[k] * er[k];
  /* virial (582-604): vrmat(j,i) += d00(j)*vr(i) */
  for (int a = 0; a < 9; a++) vr[a] = 0.0;
  for (long k = 0; k < np; k++) {
    const int ii = pi[k] * ns + c, jj = pj[k] * ns + c;
    double d00[3] = {x[ii] - x[jj] - ox[k], y[ii] - y[jj] - oy[k], z[ii] - z[jj] - oz[k]};
    double v[3] = {0, 0, 0};
    for (int l = 0; l < ns; l++)
      for (int a = 0; a < 3; a++) v[a] = v[a] + FI(k, l, a);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) vr[3 * j + i] = vr[3 * j + i] + d00[j] * v[i];
  }
#undef FI
#undef FJ
  free(fi);
  free(fj);
}

void g6o_energy_std(int model, long np, const int *pi, const int *pj,
                    const double *ox, const double *oy, const double *oz,
                    const double *er, const double *x, const double *y,
                    const double *z, double *ep) {
  g6o_model m;
  g6o_model_get(model, &m);
  const int ns = m.nsite;
  stdterm t[G6O_MAXSITE * G6O_MAXSITE];
  const int nt = std_terms(&m, t);
  for (long k = 0; k < np; k++) ep[k] = 0.0;
  for (int u = 0; u < nt; u++) {
    const stdterm s = t[u];
    for (long k = 0; k < np; k++) {
      const int ii = pi[k] * ns + s.a, jj = pj[k] * ns + s.b;
      double dx = x[ii] - x[jj], dy = y[ii] - y[jj], dz = z[ii] - z[jj];
      dx = dx - ox[k]; dy = dy - oy[k]; dz = dz - oz[k];
      double dd = 1.0 / (dx * dx + dy * dy + dz * dz), e0;
      if (s.kind == 1) {
        double ri = sqrt(dd);
        e0 = s.qq * ri;
        double d6 = dd * dd * dd, d12 = d6 * d6;
        e0 = e0 + s.aa * d12 + s.bb * d6;
      } else if (s.kind == 2) {
        e0 = s.qq * sqrt(dd);
      } else {
        double d6 = dd * dd * dd, d12 = d6 * d6;
        e0 = s.aa * d12 + s.bb * d6;
      }
      ep[k] = ep[k] + e0;
    }
  }
  for (long k = 0; k < np; k++) ep[k] = ep[k] * er[k];
}

/* ---- ST2 custom kernel with the short-range switch (ST2Module.F90:112-897) */
void g6o_force_st2(long np, const int *pi, const int *pj, const double *ox,
                   const double *oy, const double *oz, const double *er,
                   const double *fr, const double *x, const double *y,
                   const double *z, double *fx, double *fy, double *fz,
                   double *ep1, double *vr) {
  double K[12];
  g6o_constants(K);
  const double QMM = K[5], QHM = K[6], QHH = K[7], AA = K[8], BB = K[9];
  const double rl = K[10], ru = K[11];
  /* Coulomb terms: sites 2..5 (1-based) = H,H,M,M; 16 pairs */
  for (int a = 0; a < 9; a++) vr[a] = 0.0;
  for (long k = 0; k < np; k++) {
    const int ii = pi[k] * 6, jj = pj[k] * 6;
    double dx00 = x[ii + 5] - x[jj + 5], dy00 = y[ii + 5] - y[jj + 5],
           dz00 = z[ii + 5] - z[jj + 5];
    dx00 = dx00 - ox[k]; dy00 = dy00 - oy[k]; dz00 = dz00 - oz[k];
    const double hh = er[k], dh = fr[k];
    double d12 = dx00 * dx00 + dy00 * dy00 + dz00 * dz00, s12, ds12;
    if (d12 < ru * ru) {
      if (d12 < rl * rl) { s12 = 0.0; ds12 = 0.0; }
      else {
        double r12 = sqrt(d12);
        s12 = (r12 - rl) * (r12 - rl) * (3.0 * ru - rl - 2.0 * r12) / ipow(ru - rl, 3);
        ds12 = 6.0 * (r12 - rl) * (ru - r12) / ipow(ru - rl, 3);
        ds12 = -ds12 / r12;
      }
    } else { s12 = 1.0; ds12 = 0.0; }
    const double dhc = dh * s12 + hh * ds12, hhc = hh * s12;
    double dhcx = dhc * dx00, dhcy = dhc * dy00, dhcz = dhc * dz00;
    double dhx = dh * dx00, dhy = dh * dy00, dhz = dh * dz00;
    const double cx = -ox[k], cy = -oy[k], cz = -oz[k];
    double fi[6][3] = {{0}}, fj[6][3] = {{0}};
    double ep = 0.0, eplj;
    for (int a = 1; a <= 4; a++)
      for (int b = 1; b <= 4; b++) {
        const double qq = (a <= 2) == (b <= 2) ? (a <= 2 ? QHH : QMM) : QHM;
        double dx = x[ii + a] - x[jj + b] + cx, dy = y[ii + a] - y[jj + b] + cy,
               dz = z[ii + a] - z[jj + b] + cz;
        double dd = 1.0 / (dx * dx + dy * dy + dz * dz), ri = sqrt(dd);
        double e0 = qq * ri, f0 = qq * ri * dd;
        ep = ep + e0;
        fi[a][0] += dx * f0; fi[a][1] += dy * f0; fi[a][2] += dz * f0;
        fj[b][0] += dx * f0; fj[b][1] += dy * f0; fj[b][2] += dz * f0;
      }
    {
      double dx = x[ii] - x[jj] + cx, dy = y[ii] - y[jj] + cy, dz = z[ii] - z[jj] + cz;
      double dd = 1.0 / (dx * dx + dy * dy + dz * dz);
      double d6 = dd * dd * dd, dl12 = d6 * d6;
      eplj = AA * dl12 + BB * d6;
      double f0 = dd * (AA * 12.0 * dl12 + BB * 6.0 * d6);
      fi[0][0] += dx * f0; fi[0][1] += dy * f0; fi[0][2] += dz * f0;
      fj[0][0] += dx * f0; fj[0][1] += dy * f0; fj[0][2] += dz * f0;
    }
    for (int s = 0; s < 6; s++) {
      const double sc = (s == 0 || s == 5) ? hh : hhc;
      for (int a = 0; a < 3; a++) { fi[s][a] *= sc; fj[s][a] *= sc; }
    }
    dhx *= eplj; dhy *= eplj; dhz *= eplj;
    dhcx *= ep; dhcy *= ep; dhcz *= ep;
    fi[5][0] += dhx + dhcx; fi[5][1] += dhy + dhcy; fi[5][2] += dhz + dhcz;
    fj[5][0] += dhx + dhcx; fj[5][1] += dhy + dhcy; fj[5][2] += dhz + dhcz;
    ep1[k] = ep * hhc + eplj * hh;
    double v[3];
    for (int a = 0; a < 3; a++)
      v[a] = fi[0][a] + fi[1][a] + fi[2][a] + fi[3][a] + fi[4][a] + fi[5][a];
    const double d00[3] = {dx00, dy00, dz00};
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) vr[3 * a + b] += d00[a] * v[b];
    for (int s = 0; s < 6; s++) {
      fx[ii + s] += fi[s][0]; fy[ii + s] += fi[s][1]; fz[ii + s] += fi[s][2];
      fx[jj + s] -= fj[s][0]; fy[jj + s] -= fj[s][1]; fz[jj + s] -= fj[s][2];
    }
  }
}

/* ---- rigid bodies (RigidModule.F90) ------------------------------------- */
/* rigid_setmatrix (62-83) + rigid_setsiteposition0 serial path (110-128).
 * tmat[9n] row-major t11..t33; intra[3*nsite*n] per molecule per site.      */
void g6o_setsites(int model, int n, const double *com, const double *quat,
                  double *x, double *y, double *z, double *tmat, double *intra) {
  g6o_model m;
  g6o_model_get(model, &m);
  const int ns = m.nsite;
  for (int i = 0; i < n; i++) {
    const double qa = quat[0 * n + i], qb = quat[1 * n + i], qc = quat[2 * n + i],
                 qd = quat[3 * n + i];
    double *t = tmat + 9 * i;
    t[0] = (qa * qa + qb * qb - (qc * qc + qd * qd));
    t[1] = -2.0 * (qa * qd + qb * qc);
    t[2] = 2.0 * (qb * qd - qa * qc);
    t[3] = 2.0 * (qa * qd - qb * qc);
    t[4] = qa * qa + qc * qc - (qb * qb + qd * qd);
    t[5] = -2.0 * (qa * qb + qc * qd);
    t[6] = 2.0 * (qa * qc + qb * qd);
    t[7] = 2.0 * (qa * qb - qc * qd);
    t[8] = qa * qa + qd * qd - (qb * qb + qc * qc);
    for (int s = 0; s < ns; s++) {
      double *r = intra + 3 * (i * ns + s);
      r[0] = t[0] * m.bx[s] + t[1] * m.by[s] + t[2] * m.bz[s];
      r[1] = t[3] * m.bx[s] + t[4] * m.by[s] + t[5] * m.bz[s];
      r[2] = t[6] * m.bx[s] + t[7] * m.by[s] + t[8] * m.bz[s];
      x[i * ns + s] = com[3 * i] + r[0];
      y[i * ns + s] = com[3 * i + 1] + r[1];
      z[i * ns + s] = com[3 * i + 2] + r[2];
    }
  }
}

/* rigid_collectforce serial path (275-317); torque returned in body frame */
void g6o_collect(int model, int n, const double *fx, const double *fy,
                 const double *fz, const double *tmat, const double *intra,
                 double *force, double *torque) {
  g6o_model m;
  g6o_model_get(model, &m);
  const int ns = m.nsite;
  for (int i = 0; i < n; i++) {
    double F[3] = {0, 0, 0}, tq[3] = {0, 0, 0};
    for (int s = 0; s < ns; s++) {
      const int k = i * ns + s;
      const double *r = intra + 3 * k;
      F[0] = F[0] + fx[k]; F[1] = F[1] + fy[k]; F[2] = F[2] + fz[k];
      tq[0] = tq[0] + r[1] * fz[k] - r[2] * fy[k];
      tq[1] = tq[1] + r[2] * fx[k] - r[0] * fz[k];
      tq[2] = tq[2] + r[0] * fy[k] - r[1] * fx[k];
    }
    const double *t = tmat + 9 * i;
    force[3 * i] = F[0]; force[3 * i + 1] = F[1]; force[3 * i + 2] = F[2];
    torque[3 * i] = t[0] * tq[0] + t[3] * tq[1] + t[6] * tq[2];
    torque[3 * i + 1] = t[1] * tq[0] + t[4] * tq[1] + t[7] * tq[2];
    torque[3 * i + 2] = t[2] * tq[0] + t[5] * tq[1] + t[8] * tq[2];
  }
}

#define DX(s, a, i) dx[((s) * 3 + (a)) * n + (i)]
#define W(s, a, i) w[((s) * 3 + (a)) * n + (i)]
#define Q(s, a, i) quat[((s) * 4 + (a)) * n + (i)]
/* rigid_predict (480-559) */
void g6o_predict(int n, double *com, double *dx, double *w, double *quat) {
  for (int i = 0; i < n; i++) {
    for (int a = 0; a < 3; a++) {
      com[3 * i + a] = com[3 * i + a] + DX(0, a, i) + DX(1, a, i) + DX(2, a, i) +
                       DX(3, a, i) + DX(4, a, i);
      DX(0, a, i) = DX(0, a, i) + 2.0 * DX(1, a, i) + 3.0 * DX(2, a, i) +
                    4.0 * DX(3, a, i) + 5.0 * DX(4, a, i);
      DX(1, a, i) = DX(1, a, i) + 3.0 * DX(2, a, i) + 6.0 * DX(3, a, i) + 10.0 * DX(4, a, i);
      DX(2, a, i) = DX(2, a, i) + 4.0 * DX(3, a, i) + 10.0 * DX(4, a, i);
      DX(3, a, i) = DX(3, a, i) + 5.0 * DX(4, a, i);
    }
    for (int a = 0; a < 3; a++) {
      W(0, a, i) = W(0, a, i) + W(1, a, i) + W(2, a, i) + W(3, a, i) + W(4, a, i);
      W(1, a, i) = W(1, a, i) + 2.0 * W(2, a, i) + 3.0 * W(3, a, i) + 4.0 * W(4, a, i);
      W(2, a, i) = W(2, a, i) + 3.0 * W(3, a, i) + 6.0 * W(4, a, i);
      W(3, a, i) = W(3, a, i) + 4.0 * W(4, a, i);
    }
    for (int a = 0; a < 4; a++) {
      Q(0, a, i) = Q(0, a, i) + Q(1, a, i) + Q(2, a, i) + Q(3, a, i) + Q(4, a, i);
      Q(1, a, i) = Q(1, a, i) + 2.0 * Q(2, a, i) + 3.0 * Q(3, a, i) + 4.0 * Q(4, a, i);
    }
    const double re = 1.0 / sqrt(Q(0, 0, i) * Q(0, 0, i) + Q(0, 1, i) * Q(0, 1, i) +
                                 Q(0, 2, i) * Q(0, 2, i) + Q(0, 3, i) * Q(0, 3, i));
    for (int a = 0; a < 4; a++) Q(0, a, i) = Q(0, a, i) * re;
  }
}

/* rigid_correct (1255-1487), Nose drag only (no Andersen) */
void g6o_correct(int model, int n, double dt, double zeta0, const double *force,
                 const double *torque, double *com, double *dx, double *w,
                 double *quat) {
  g6o_correct_npt(model, n, dt, zeta0, 0.0, force, torque, com, dx, w, quat);
}

/* the same with the Andersen dilation term acoeff = v1 / (2*3*v0) added to
 * the translational drag (RigidModule.F90:1290-1305, orthorhombic) */
void g6o_correct_npt(int model, int n, double dt, double zeta0, double acoeff,
                     const double *force, const double *torque, double *com, double *dx,
                     double *w, double *quat) {
  g6o_model m;
  g6o_model_get(model, &m);
  const double massi = 1.0 / m.mass, Ixxi = 1.0 / m.Ixx, Iyyi = 1.0 / m.Iyy,
               Izzi = 1.0 / m.Izz;
  const double nc = zeta0 * dt, exr = nc;
  const double ext = nc * 0.5 + acoeff;
  const double h = dt * 0.5, tf = dt * dt * 0.5 * massi;
  for (int i = 0; i < n; i++) {
    for (int a = 0; a < 3; a++) {
      double nf = force[3 * i + a] * tf;
      nf = nf - DX(0, a, i) * ext;
      const double d = nf - DX(1, a, i);
      com[3 * i + a] = com[3 * i + a] + Gear5_0 * d;
      DX(0, a, i) = DX(0, a, i) + Gear5_1 * d;
      DX(1, a, i) = nf;
      DX(2, a, i) = DX(2, a, i) + Gear5_3 * d;
      DX(3, a, i) = DX(3, a, i) + Gear5_4 * d;
      DX(4, a, i) = DX(4, a, i) + Gear5_5 * d;
    }
    double wx = W(0, 0, i), wy = W(0, 1, i), wz = W(0, 2, i);
    double qa = Q(0, 0, i), qb = Q(0, 1, i), qc = Q(0, 2, i), qd = Q(0, 3, i);
    const double nwa = (-wx * qb + wy * qc - wz * qd) * dt * 0.5;
    const double nwb = (wx * qa - wy * qd - wz * qc) * dt * 0.5;
    const double nwc = (-wx * qd - wy * qa + wz * qb) * dt * 0.5;
    const double nwd = (wx * qc + wy * qb + wz * qa) * dt * 0.5;
    const double dwa = nwa - Q(1, 0, i), dwb = nwb - Q(1, 1, i), dwc = nwc - Q(1, 2, i),
                 dwd = nwd - Q(1, 3, i);
    Q(0, 0, i) = Q(0, 0, i) + Gear4_0 * dwa;
    Q(0, 1, i) = Q(0, 1, i) + Gear4_0 * dwb;
    Q(0, 2, i) = Q(0, 2, i) + Gear4_0 * dwc;
    Q(0, 3, i) = Q(0, 3, i) + Gear4_0 * dwd;
    double nf[3];
    nf[0] = (torque[3 * i] + (m.Iyy - m.Izz) * wy * wz) * dt * Ixxi;
    nf[1] = (torque[3 * i + 1] + (m.Izz - m.Ixx) * wz * wx) * dt * Iyyi;
    nf[2] = (torque[3 * i + 2] + (m.Ixx - m.Iyy) * wx * wy) * dt * Izzi;
    for (int a = 0; a < 3; a++) {
      nf[a] = nf[a] - W(0, a, i) * exr;
      const double d = nf[a] - W(1, a, i);
      W(0, a, i) = W(0, a, i) + Gear4_0 * d;
      W(1, a, i) = nf[a];
      W(2, a, i) = W(2, a, i) + Gear4_2 * d;
      W(3, a, i) = W(3, a, i) + Gear4_3 * d;
      W(4, a, i) = W(4, a, i) + Gear4_4 * d;
    }
    /* quaternion derivatives 2..5 from the corrected omega derivatives */
    double w1x = W(0, 0, i), w1y = W(0, 1, i), w1z = W(0, 2, i);
    double w2x = W(1, 0, i), w2y = W(1, 1, i), w2z = W(1, 2, i);
    double w3x = W(2, 0, i), w3y = W(2, 1, i), w3z = W(2, 2, i);
    double w4x = W(3, 0, i), w4y = W(3, 1, i), w4z = W(3, 2, i);
    qa = Q(0, 0, i); qb = Q(0, 1, i); qc = Q(0, 2, i); qd = Q(0, 3, i);
    Q(1, 2, i) = (-w1x * qd - w1y * qa + w1z * qb) * dt * 0.5;
    Q(1, 1, i) = (w1x * qa - w1y * qd - w1z * qc) * dt * 0.5;
    Q(1, 3, i) = (w1x * qc + w1y * qb + w1z * qa) * dt * 0.5;
    Q(1, 0, i) = (-w1x * qb + w1y * qc - w1z * qd) * dt * 0.5;
    const double pa = Q(1, 0, i), pb = Q(1, 1, i), pc = Q(1, 2, i), pd = Q(1, 3, i);
    Q(2, 2, i) = (-w2x * qd - w2y * qa + w2z * qb - w1x * pd - w1y * pa + w1z * pb) * dt * 0.25;
    Q(2, 1, i) = (w2x * qa - w2y * qd - w2z * qc + w1x * pa - w1y * pd - w1z * pc) * dt * 0.25;
    Q(2, 3, i) = (w2x * qc + w2y * qb + w2z * qa + w1x * pc + w1y * pb + w1z * pa) * dt * 0.25;
    Q(2, 0, i) = (-w2x * qb + w2y * qc - w2z * qd - w1x * pb + w1y * pc - w1z * pd) * dt * 0.25;
    const double ra = Q(2, 0, i), rb = Q(2, 1, i), rc_ = Q(2, 2, i), rd = Q(2, 3, i);
    Q(3, 2, i) = (-w3x * qd - w3y * qa + w3z * qb - w1x * rd - w1y * ra + w1z * rb -
                  w2x * pd - w2y * pa + w2z * pb) * dt * (1.0 / 6.0);
    Q(3, 1, i) = (w3x * qa - w3y * qd - w3z * qc + w1x * ra - w1y * rd - w1z * rc_ +
                  w2x * pa - w2y * pd - w2z * pc) * dt * (1.0 / 6.0);
    Q(3, 3, i) = (w3x * qc + w3y * qb + w3z * qa + w1x * rc_ + w1y * rb + w1z * ra +
                  w2x * pc + w2y * pb + w2z * pa) * dt * (1.0 / 6.0);
    Q(3, 0, i) = (-w3x * qb + w3y * qc - w3z * qd - w1x * rb + w1y * rc_ - w1z * rd -
                  w2x * pb + w2y * pc - w2z * pd) * dt * (1.0 / 6.0);
    const double sa = Q(3, 0, i), sb = Q(3, 1, i), sc = Q(3, 2, i), sd = Q(3, 3, i);
    Q(4, 2, i) = (-w4x * qd - w4y * qa + w4z * qb - w3x * pd - w3y * pa + w3z * pb -
                  w2x * rd - w2y * ra + w2z * rb - w1x * sd - w1y * sa + w1z * sb) * dt * (1.0 / 8.0);
    Q(4, 1, i) = (w4x * qa - w4y * qd - w4z * qc + w3x * pa - w3y * pd - w3z * pc +
                  w2x * ra - w2y * rd - w2z * rc_ + w1x * sa - w1y * sd - w1z * sc) * dt * (1.0 / 8.0);
    Q(4, 3, i) = (w4x * qc + w4y * qb + w4z * qa + w3x * pc + w3y * pb + w3z * pa +
                  w2x * rc_ + w2y * rb + w2z * ra + w1x * sc + w1y * sb + w1z * sa) * dt * (1.0 / 8.0);
    Q(4, 0, i) = (-w4x * qb + w4y * qc - w4z * qd - w3x * pb + w3y * pc - w3z * pd -
                  w2x * rb + w2y * rc_ - w2z * rd - w1x * sb + w1y * sc - w1z * sd) * dt * (1.0 / 8.0);
  }
  (void)h;
}

/* Rigid_Ek (576-597) */
void g6o_ek(int model, int n, double dt, const double *dx, const double *w,
            double *ekt, double *ekr) {
  g6o_model m;
  g6o_model_get(model, &m);
  double t = 0, rx = 0, ry = 0, rz = 0;
  for (int i = 0; i < n; i++) {
    t = t + DX(0, 0, i) * DX(0, 0, i) + DX(0, 1, i) * DX(0, 1, i) + DX(0, 2, i) * DX(0, 2, i);
    rx = rx + W(0, 0, i) * W(0, 0, i);
    ry = ry + W(0, 1, i) * W(0, 1, i);
    rz = rz + W(0, 2, i) * W(0, 2, i);
  }
  *ekr = (rx * m.Ixx + ry * m.Iyy + rz * m.Izz) * 0.5;
  *ekt = t * m.mass * 0.5 * (1.0 / (dt * dt));
}
#undef DX
#undef W
#undef Q

/* Nose_Predict / Nose_Correct (NoseModule.F90:43-87) */
void g6o_nose_predict(double *s) {
  s[3] = s[3] + s[4] + s[5] + s[6] + s[7];
  s[4] = s[4] + 2.0 * s[5] + 3.0 * s[6] + 4.0 * s[7];
  s[5] = s[5] + 3.0 * s[6] + 6.0 * s[7];
  s[6] = s[6] + 4.0 * s[7];
}
void g6o_nose_correct(double *s, double temp) {
  const double zetaf = (temp - s[0]) / s[1] - s[2] * s[3];
  const double dz2 = zetaf - s[4];
  s[3] = s[3] + Gear4_0 * dz2;
  s[4] = zetaf;
  s[5] = s[5] + Gear4_2 * dz2;
  s[6] = s[6] + Gear4_3 * dz2;
  s[7] = s[7] + Gear4_4 * dz2;
  s[8] = s[8] + s[3];
}

/* ---- MD step in Mixture.F90 order (1789-2286) --------------------------- */
/* Andersen_Rescale, orthorhombic (AndersenModule.F90:65-82, BoxModule.F90:
 * 176-183): the cube takes the volume v0 */
static void andersen_rescale(const double *a, double *L) {
  const double r2 = pow(a[2] / ((*L * *L) * *L), 1.0 / 3.0);
  *L = *L * r2;
}

/* Andersen_Predict (AndersenModule.F90:84-93); a = pex, mass, v0..v5 */
void g6o_andersen_predict(double *a, double *L) {
  a[2] = a[2] + a[3] + a[4] + a[5] + a[6] + a[7];
  a[3] = a[3] + 2.0 * a[4] + 3.0 * a[5] + 4.0 * a[6] + 5.0 * a[7];
  a[4] = a[4] + 3.0 * a[5] + 6.0 * a[6] + 10.0 * a[7];
  a[5] = a[5] + 4.0 * a[6] + 10.0 * a[7];
  a[6] = a[6] + 5.0 * a[7];
  andersen_rescale(a, L);
}

/* Andersen_Correct (AndersenModule.F90:95-110), pin in atm */
void g6o_andersen_correct(double *a, double *L, double pin) {
  const double volp2 = (pin - a[0]) / a[1] * 0.5;
  const double cvol = volp2 - a[4];
  a[2] = a[2] + Gear5_0 * cvol;
  a[3] = a[3] + Gear5_1 * cvol;
  a[4] = volp2;
  a[5] = a[5] + Gear5_3 * cvol;
  a[6] = a[6] + Gear5_4 * cvol;
  a[7] = a[7] + Gear5_5 * cvol;
  andersen_rescale(a, L);
}

long g6o_md_step(g6o_md *s) {
  g6o_model m;
  g6o_model_get(s->model, &m);
  const int n = s->n, ns = m.nsite;
  g6o_predict(n, s->com, s->dx, s->w, s->quat);
  if (s->nose_active) g6o_nose_predict(s->nose);
  if (s->andersen) g6o_andersen_predict(s->anders, &s->L);
  double *x = malloc(sizeof(double) * n * ns * 6), *y = x + n * ns, *z = y + n * ns;
  double *fx = z + n * ns, *fy = fx + n * ns, *fz = fy + n * ns;
  double *tmat = malloc(sizeof(double) * 9 * n), *intra = malloc(sizeof(double) * 3 * ns * n);
  g6o_setsites(s->model, n, s->com, s->quat, x, y, z, tmat, intra);
  const long cap = (long)n * 800;
  int *pi = malloc(sizeof(int) * cap), *pj = malloc(sizeof(int) * cap);
  double *o = malloc(sizeof(double) * cap * 6);
  double *ox = o, *oy = o + cap, *oz = o + 2 * cap, *er = o + 3 * cap, *fr = o + 4 * cap,
         *ep1 = o + 5 * cap;
  long np0;
  long np = g6o_pairlist(n, s->com, s->L, s->rc, s->req, cap, &np0, pi, pj, ox, oy, oz);
  if (np >= 0) {
    memset(fx, 0, sizeof(double) * 3 * n * ns);
    g6o_cutoff_dimmer(np, pi, pj, s->com, ox, oy, oz, s->rc - s->taper, s->rc, er, fr);
    if (s->model == G6O_TIP4P && s->fast)
      g6o_force_tip4p(np, pi, pj, ox, oy, oz, er, fr, x, y, z, fx, fy, fz, ep1, s->vir);
    else if (s->model == G6O_ST2 && s->fast)
      g6o_force_st2(np, pi, pj, ox, oy, oz, er, fr, x, y, z, fx, fy, fz, ep1, s->vir);
    else
      g6o_force_std(s->model, np, pi, pj, ox, oy, oz, er, fr, x, y, z, fx, fy, fz, ep1, s->vir);
    double ep = 0.0;
    for (long k = 0; k < np; k++) ep = ep + ep1[k];
    s->ep = ep;
    double *F = malloc(sizeof(double) * 6 * n), *T = F + 3 * n;
    g6o_collect(s->model, n, fx, fy, fz, tmat, intra, F, T);
    const double acoeff = s->andersen ? s->anders[3] / (2.0 * 3.0 * s->anders[2]) : 0.0;
    g6o_correct_npt(s->model, n, s->dt, s->nose_active ? s->nose[3] : 0.0, acoeff, F, T,
                    s->com, s->dx, s->w, s->quat);
    free(F);
    g6o_ek(s->model, n, s->dt, s->dx, s->w, &s->ekt, &s->ekr);
    s->temperature = I2J * 2.0 * (s->ekt + s->ekr) / (rgas() * (double)(6 * n));
    /* Rigid_PressureTensor (RigidModule.F90:600-619) + the pair virial, and
     * the pressure in atm (Mixture.F90:2236-2255) */
    double kt[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++)
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) kt[3 * a + b] = kt[3 * a + b] + s->dx[a * n + i] * s->dx[b * n + i];
    for (int q = 0; q < 9; q++) kt[q] = kt[q] * m.mass * (1.0 / (s->dt * s->dt));
    const double vp = (kt[0] + s->vir[0]) + (kt[4] + s->vir[4]) + (kt[8] + s->vir[8]);
    s->pressure = (vp / 3.0) * I2J * (1.0 / 101326.0) /
                  (((s->L * s->L) * s->L) * 6.0225e23 * (1e-10 * 1e-10 * 1e-10));
    if (s->nose_active) g6o_nose_correct(s->nose, s->temperature);
    if (s->andersen) g6o_andersen_correct(s->anders, &s->L, s->pressure);
    s->volume = (s->L * s->L) * s->L;
    s->andersen_ep = s->andersen ? s->anders[0] * s->anders[2] * 101326.0 * 6.0225e23 *
                                       (1e-10 * 1e-10 * 1e-10)
                                 : 0.0;
    s->andersen_ek = s->andersen ? s->anders[3] * s->anders[3] * s->anders[1] * 0.5 *
                                       101326.0 * 6.0225e23 * (1e-10 * 1e-10 * 1e-10)
                                 : 0.0;
  }
  s->npair = np;
  free(x); free(tmat); free(intra); free(pi); free(pj); free(o);
  return np;
}

double g6o_md_hamiltonian(const g6o_md *s) {
  const double dof = 6.0 * s->n;
  double h = (s->ep + s->ekt + s->ekr) * I2J;
  if (s->nose_active) {
    const double *z = s->nose;
    const double nose_ek = (z[3] * z[3] * z[1] * s->dt * dof * 0.5) * rgas();
    const double nose_ep = (dof * z[0] * (z[8] - 0.5 * z[3]) * s->dt) * rgas();
    h = h + nose_ep + nose_ek;
  }
  if (s->andersen) h = h + s->andersen_ep + s->andersen_ek;
  return h;
}

/* ---- MT19937 (Matsumoto & Nishimura 1998, init_genrand of 2002) --------- */
static uint32_t mt[624];
static int mti = 625;
void g6o_mt_seed(uint32_t seed) {
  mt[0] = seed;
  for (int i = 1; i < 624; i++)
    mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
  mti = 624;
}
uint32_t g6o_mt_next(void) {
  if (mti >= 624) {
    if (mti == 625) g6o_mt_seed(5489u);
    for (int k = 0; k < 624; k++) {
      uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
      mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    mti = 0;
  }
  uint32_t y = mt[mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}
double g6o_random_getnext(void) {
  const double x = (double)(g6o_mt_next() >> 1) * (1.0 / 2147483648.0);
  return 1.0 - x;
}

/* Random_Normal (RandomModule.F90:28-36) */
static double random_normal(double avg, double var, double r1, double r2) {
  double r = sqrt(-2.0 * log(r1));
  double theta = 2.0 * PI_T * r2;
  return r * sin(theta) * sqrt(var) + avg;
}

void g6o_vreset(int model, int n, double dt, double temp, uint32_t seed,
                double *dx, double *w) {
  g6o_model m;
  g6o_model_get(model, &m);
  const double massi = 1.0 / m.mass, iin[3] = {1.0 / m.Ixx, 1.0 / m.Iyy, 1.0 / m.Izz};
  g6o_mt_seed(seed);
  for (int i = 0; i < n; i++) {
    for (int a = 0; a < 3; a++) {
      double r1 = g6o_random_getnext(), r2 = g6o_random_getnext();
      dx[a * n + i] = random_normal(0.0, temp * rgas() * J2I() * massi, r1, r2) * dt;
    }
    for (int a = 0; a < 3; a++) {
      double r1 = g6o_random_getnext(), r2 = g6o_random_getnext();
      w[a * n + i] = random_normal(0.0, temp * rgas() * J2I() * iin[a], r1, r2);
    }
  }
}

/* ---- Monte Carlo: PairInfo and the mc2 trial loop ----------------------- */
/* StdPairPotential of one pair from site arrays (StdInteraction.F90:633-762):
 * site a of molecule i (xi..) with site b of molecule j, shift o, times er */
static double std_pair(const stdterm *t, int nt, int ns, const double *x, const double *y,
                       const double *z, int i, const double *xt, const double *yt,
                       const double *zt, double ox, double oy, double oz, double er) {
  double ep = 0.0;
  for (int u = 0; u < nt; u++) {
    const stdterm s = t[u];
    double dx = x[i * ns + s.a] - xt[s.b], dy = y[i * ns + s.a] - yt[s.b],
           dz = z[i * ns + s.a] - zt[s.b];
    dx = dx - ox; dy = dy - oy; dz = dz - oz;
    double dd = 1.0 / (dx * dx + dy * dy + dz * dz), e0;
    if (s.kind == 1) {
      double ri = sqrt(dd);
      e0 = s.qq * ri;
      double d6 = dd * dd * dd, d12 = d6 * d6;
      e0 = e0 + s.aa * d12 + s.bb * d6;
    } else if (s.kind == 2) {
      e0 = s.qq * sqrt(dd);
    } else {
      double d6 = dd * dd * dd, d12 = d6 * d6;
      e0 = s.aa * d12 + s.bb * d6;
    }
    ep = ep + e0;
  }
  return ep * er;
}

/* PairInfo as mc2 builds it (mc2.F90:303-340): interaction_alltoall
 * (InteractionModule.F90:157-196), Interaction_Compress with rc, CutOff_Dimmer2
 * and pairenergy (PairInfo.F90:12-46).  com[n][3]; quat[4][n] (order 1);
 * ep[n*n] symmetric, zero off the list; returns epsum in list order. */
double g6o_mc_pairinfo(int model, int n, double L, double rc, double taper,
                       const double *com, const double *quat, double *ep) {
  g6o_model m;
  g6o_model_get(model, &m);
  const int ns = m.nsite;
  double *x = malloc(sizeof(double) * n * ns), *y = malloc(sizeof(double) * n * ns),
         *z = malloc(sizeof(double) * n * ns), *t = malloc(sizeof(double) * 9 * n),
         *intra = malloc(sizeof(double) * 3 * ns * n);
  g6o_setsites(model, n, com, quat, x, y, z, t, intra);
  stdterm st[G6O_MAXSITE * G6O_MAXSITE];
  const int nt = std_terms(&m, st);
  for (long k = 0; k < (long)n * n; k++) ep[k] = 0.0;
  double epsum = 0.0;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      double d[3], o[3];
      for (int a = 0; a < 3; a++) {
        d[a] = com[3 * i + a] - com[3 * j + a];
        o[a] = round(d[a] * (1.0 / L)) * L;
        d[a] = d[a] - o[a];
      }
      const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (!(dd < rc * rc)) continue;
      double hh, dh;
      g6o_cutoff_func(rc - taper, rc, dd, &hh, &dh);
      const double e = std_pair(st, nt, ns, x, y, z, i, x + j * ns, y + j * ns, z + j * ns,
                                o[0], o[1], o[2], hh);
      ep[(long)i * n + j] = e;
      ep[(long)j * n + i] = e;
      epsum = epsum + e;
    }
  free(x); free(y); free(z); free(t); free(intra);
  return epsum;
}

/* The mc2 single-molecule trial loop (mc2.F90:540-862, DEBUG20050322a) on
 * the Random_GetNext stream u starting at *pos; the two draws of each
 * random_normal are taken in argument order.  com[n][3], quat[4][n],
 * delta[n], rtr[n], ep[n*n] (PairInfo), *epsum are updated in place;
 * returns the number of accepted trials. */
typedef struct umbctx umbctx;
static int umb_prepare(umbctx *U, int n, double L, double rc, const double *com, double epsum);
static double umb_trial(umbctx *U, int target, double dx, double dy, double dz, double epsum,
                        double deltae);
static void umb_accept(umbctx *U);
static void umb_reject(umbctx *U);

static long mc_trials(int model, int n, double L, double rc, double taper, double temp,
                      double *com, double *quat, double *delta, double *rtr, double *ep,
                      double *epsum, const double *u, long *pos, int ntrial, umbctx *U) {
  const double INCREMENT = (double)1.05f, I2J_ = 10.0, J2K_ = 0.120273, beta = 1.0 / temp;
  g6o_model m;
  g6o_model_get(model, &m);
  const int ns = m.nsite;
  double *x = malloc(sizeof(double) * n * ns), *y = malloc(sizeof(double) * n * ns),
         *z = malloc(sizeof(double) * n * ns), *t = malloc(sizeof(double) * 9 * n),
         *intra = malloc(sizeof(double) * 3 * ns * n), *enew = malloc(sizeof(double) * n);
  char *inl = malloc(n);
  g6o_setsites(model, n, com, quat, x, y, z, t, intra);
  stdterm st[G6O_MAXSITE * G6O_MAXSITE];
  const int nt = std_terms(&m, st);
  long p = *pos, nacc = 0;
  if (U && umb_prepare(U, n, L, rc, com, *epsum)) ntrial = -1;
  for (int trial = 0; trial < ntrial; trial++) {
    int target = (int)(u[p++] * n + 1) - 1;
    if (target >= n) target = n - 1;
    const double dt = delta[target], dr = delta[target] * rtr[target];
    const double rotx = random_normal(0.0, dr * dr, u[p], u[p + 1]);
    const double roty = random_normal(0.0, dr * dr, u[p + 2], u[p + 3]);
    const double rotz = random_normal(0.0, dr * dr, u[p + 4], u[p + 5]);
    p += 6;
    const double avgrot = (fabs(rotx) + fabs(roty) + fabs(rotz)) / 3.0;
    double qa = 1.0, qb = rotx, qc = roty, qd = rotz; /* qinfrotator + qnormalize */
    {
      const double r = 1.0 / sqrt(qa * qa + qb * qb + qc * qc + qd * qd);
      qa = qa * r; qb = qb * r; qc = qc * r; qd = qd * r;
    }
    { /* qadd (QuatModule.F90:108-133) */
      const double a2 = quat[target], b2 = quat[n + target], c2 = quat[2 * n + target],
                   d2 = quat[3 * n + target];
      const double a3 = qa * a2 - qb * b2 - qc * c2 - qd * d2;
      const double b3 = qa * b2 + qb * a2 + qc * d2 - qd * c2;
      const double c3 = qa * c2 + qc * a2 - qb * d2 + qd * b2;
      const double d3 = qa * d2 + qd * a2 + qb * c2 - qc * b2;
      qa = a3; qb = b3; qc = c3; qd = d3;
      if (qa < 0) { qa = -qa; qb = -qb; qc = -qc; qd = -qd; }
    }
    const double deltax = random_normal(0.0, dt * dt, u[p], u[p + 1]);
    const double dum1 = u[p + 2], dum2 = u[p + 3];
    const double deltay = random_normal(0.0, dt * dt, dum1, dum2);
    const double deltaz = random_normal(0.0, dt * dt, u[p + 4], u[p + 5]);
    p += 6;
    const double avgtrans = (fabs(deltax) + fabs(deltay) + fabs(deltaz)) / 3.0;
    const double ct[3] = {deltax + com[3 * target], deltay + com[3 * target + 1],
                          deltaz + com[3 * target + 2]};
    double qt[4] = {qa, qb, qc, qd}, xt[G6O_MAXSITE], yt[G6O_MAXSITE], zt[G6O_MAXSITE], tt[9],
           it[3 * G6O_MAXSITE];
    g6o_setsites(model, 1, ct, qt, xt, yt, zt, tt, it);
    /* the compressed list (i, moved target) within rc + dt, ascending i */
    const double rr = (rc + dt) * (rc + dt);
    for (int i = 0; i < n; i++) {
      inl[i] = 0;
      enew[i] = 0.0;
      const double *ci = (i == target) ? ct : com + 3 * i;
      double d[3], o[3];
      for (int a = 0; a < 3; a++) {
        d[a] = ci[a] - ct[a];
        o[a] = round(d[a] * (1.0 / L)) * L;
        d[a] = d[a] - o[a];
      }
      const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (!(dd < rr)) continue;
      inl[i] = 1;
      if (i == target) continue;
      double hh, dh;
      g6o_cutoff_func(rc - taper, rc, dd, &hh, &dh);
      enew[i] = std_pair(st, nt, ns, x, y, z, i, xt, yt, zt, o[0], o[1], o[2], hh);
    }
    double deltae = 0.0;
    for (int i = 0; i < n; i++) deltae = deltae - ep[(long)i * n + target];
    for (int i = 0; i < n; i++)
      if (inl[i] && i != target) deltae = deltae + enew[i];
    /* deltaesum = deltae + the umbrellas' change (mc2.F90:684-706) */
    const double deltaesum = U ? umb_trial(U, target, deltax, deltay, deltaz, *epsum, deltae)
                               : deltae;
    int accept = 0;
    if (deltaesum < 0.0) {
      accept = 1;
    } else {
      const double ratio = exp(-(deltaesum * I2J_ * J2K_) * beta);
      if (u[p++] < ratio) accept = 1;
    }
    if (accept) {
      nacc++;
      if (U) umb_accept(U);
      for (int a = 0; a < 3; a++) com[3 * target + a] = ct[a];
      for (int a = 0; a < 4; a++) quat[a * n + target] = qt[a];
      for (int s = 0; s < ns; s++) {
        x[target * ns + s] = xt[s]; y[target * ns + s] = yt[s]; z[target * ns + s] = zt[s];
      }
      *epsum = *epsum + deltae;
      for (int i = 0; i < n; i++) {
        ep[(long)i * n + target] = 0.0;
        ep[(long)target * n + i] = 0.0;
      }
      for (int i = 0; i < n; i++)
        if (inl[i] && i != target) {
          ep[(long)i * n + target] = enew[i];
          ep[(long)target * n + i] = enew[i];
        }
      delta[target] = delta[target] * INCREMENT;
      if (avgtrans * rtr[target] < avgrot)
        rtr[target] = rtr[target] * INCREMENT;
      else
        rtr[target] = rtr[target] / INCREMENT;
    } else {
      if (U) umb_reject(U);
      delta[target] = delta[target] / INCREMENT;
      if (delta[target] < (double)0.05f) delta[target] = (double)0.05f;
    }
  }
  *pos = p;
  free(x); free(y); free(z); free(t); free(intra); free(enew); free(inl);
  return ntrial < 0 ? -1 : nacc;
}

long g6o_mc_trials(int model, int n, double L, double rc, double taper, double temp,
                   double *com, double *quat, double *delta, double *rtr, double *ep,
                   double *epsum, const double *u, long *pos, int ntrial) {
  return mc_trials(model, n, L, rc, taper, temp, com, quat, delta, rtr, ep, epsum, u, pos,
                   ntrial, NULL);
}

/* ---- NPT Monte Carlo with collective moves ------------------------------
 * No reference counterpart: the reference has single-molecule NVT moves
 * only.  The energy is PairInfo (as g6o_mc_pairinfo); the moves, the
 * proposal stream and the acceptance rule are the ones the device path
 * defines (genesis6_amd/csrc/g6_npt.hip), restated here serially. */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) */
void g6o_philox10(uint32_t *c, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1, n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

static double npt_u01(uint32_t x) { return ((double)x + 0.5) / 4294967296.0; }

typedef struct {
  int type;
  double a, x, uacc, u1;
} npt_prop;

typedef struct {
  uint32_t k0, k1;
  double pvol, dlnv, dmax, rot_e;
} npt_cfg;

/* degree-5 Taylor polynomial of exp (the volume factor of a proposal) */
static double npt_exp_small(double y) {
  return 1.0 + y * (1.0 + y * (0.5 + y * (1.0 / 6.0 + y * (1.0 / 24.0 + y * (1.0 / 120.0)))));
}

static npt_prop npt_make_prop(const npt_cfg *g, unsigned long long idx) {
  uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0u, 0x4E505400u};
  g6o_philox10(c, g->k0, g->k1);
  npt_prop p;
  p.type = npt_u01(c[0]) < g->pvol;
  p.uacc = npt_u01(c[2]);
  p.u1 = npt_u01(c[3]);
  if (p.type) {
    p.x = g->dlnv * (2.0 * npt_u01(c[1]) - 1.0);
    p.a = npt_exp_small(p.x / 3.0);
  } else {
    p.x = 0.0;
    p.a = npt_u01(c[1]);
  }
  return p;
}

/* new COM and quaternion of a block member in slot k */
/* offset of the block grid of proposal idx: uniform in one cell per axis
 * (Philox slot 0xFFFF); a block takes the molecules of one cell of the
 * Address_Assign grid shifted by it */
static void npt_offset(const npt_cfg *g, unsigned long long idx, double L, int nd, double *o) {
  uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0xFFFFu, 0x4E505400u};
  g6o_philox10(c, g->k0, g->k1);
  const double cell = L / nd;
  for (int a = 0; a < 3; a++) o[a] = npt_u01(c[a]) * cell;
}

static void npt_member(const npt_cfg *g, unsigned long long idx, int k, const double *c0,
                       const double *q0, double *c1, double *q1) {
  uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)(1 + 2 * k), 0x4E505400u};
  uint32_t e[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)(2 + 2 * k), 0x4E505400u};
  g6o_philox10(c, g->k0, g->k1);
  g6o_philox10(e, g->k0, g->k1);
  for (int a = 0; a < 3; a++) c1[a] = c0[a] + g->dmax * (2.0 * npt_u01(c[a]) - 1.0);
  const double e0 = g->rot_e * (2.0 * npt_u01(c[3]) - 1.0);
  const double e1 = g->rot_e * (2.0 * npt_u01(e[0]) - 1.0);
  const double e2 = g->rot_e * (2.0 * npt_u01(e[1]) - 1.0);
  const double r = 1.0 / sqrt(1.0 + e0 * e0 + e1 * e1 + e2 * e2);
  const double w0 = r, x0 = e0 * r, y0 = e1 * r, z0 = e2 * r;
  /* (w0, x0, y0, z0) (x) q0, then normalised */
  double w = w0 * q0[0] - x0 * q0[1] - y0 * q0[2] - z0 * q0[3];
  double x = w0 * q0[1] + x0 * q0[0] + y0 * q0[3] - z0 * q0[2];
  double y = w0 * q0[2] - x0 * q0[3] + y0 * q0[0] + z0 * q0[1];
  double z = w0 * q0[3] + x0 * q0[2] - y0 * q0[1] + z0 * q0[0];
  const double rn = 1.0 / sqrt(w * w + x * x + y * y + z * z);
  q1[0] = w * rn; q1[1] = x * rn; q1[2] = y * rn; q1[3] = z * rn;
}

/* Address_Assign cell (dnint binning of the wrapped COM) */
static int npt_cell(const double *c, double L, int nd) {
  const double invL = 1.0 / L, cellinv = 1.0 / (L / nd);
  int jc[3];
  for (int a = 0; a < 3; a++) {
    double d = c[a];
    d = d - round(d * invL) * L;
    int j = (int)round(d * cellinv);
    if (j < 0) j += nd;
    j = ((j % nd) + nd) % nd;
    jc[a] = j;
  }
  return (jc[2] * nd + jc[1]) * nd + jc[0];
}

/* sites of one molecule (rigid_setmatrix + site placement) */
static void npt_place(const g6o_model *m, const double *q, const double *c, double *x, double *y,
                      double *z) {
  const double qa = q[0], qb = q[1], qc = q[2], qd = q[3];
  const double t0 = (qa * qa + qb * qb - (qc * qc + qd * qd)), t1 = -2.0 * (qa * qd + qb * qc),
               t2 = 2.0 * (qb * qd - qa * qc), t3 = 2.0 * (qa * qd - qb * qc),
               t4 = qa * qa + qc * qc - (qb * qb + qd * qd), t5 = -2.0 * (qa * qb + qc * qd),
               t6 = 2.0 * (qa * qc + qb * qd), t7 = 2.0 * (qa * qb - qc * qd),
               t8 = qa * qa + qd * qd - (qb * qb + qc * qc);
  for (int s = 0; s < m->nsite; s++) {
    x[s] = c[0] + (t0 * m->bx[s] + t1 * m->by[s] + t2 * m->bz[s]);
    y[s] = c[1] + (t3 * m->bx[s] + t4 * m->by[s] + t5 * m->bz[s]);
    z[s] = c[2] + (t6 * m->bx[s] + t7 * m->by[s] + t8 * m->bz[s]);
  }
}

/* PairInfo pair energy, lower index first (sites x..[i*ns + s], COMs com[3i + a]) */
static double npt_pair(const g6o_model *m, const stdterm *st, int nt, const double *com,
                       const double *x, const double *y, const double *z, int i, int j,
                       double L, double rc, double taper) {
  const int lo = i < j ? i : j, hi = i < j ? j : i, ns = m->nsite;
  double d[3], o[3];
  for (int a = 0; a < 3; a++) {
    d[a] = com[3 * lo + a] - com[3 * hi + a];
    o[a] = round(d[a] * (1.0 / L)) * L;
    d[a] = d[a] - o[a];
  }
  const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (!(dd < rc * rc)) return 0.0;
  double hh, dh;
  g6o_cutoff_func(rc - taper, rc, dd, &hh, &dh);
  return std_pair(st, nt, ns, x, y, z, lo, x + hi * ns, y + hi * ns, z + hi * ns, o[0], o[1],
                  o[2], hh);
}

/* nmoves plain NPT moves from proposal index *index.  com[n][3], quat[4][n]
 * (order 1), *L and *U (the PairInfo sum) are updated; accepted[nmoves];
 * counts[5] += proposals, accepted, volume proposals, volume accepted,
 * void proposals (rejected untested: a block of more than 96 molecules, a
 * block move taking a member out of its cell, a box edge below 2 rc). */
void g6o_npt_run(int model, int n, double rc, double taper, double voxa, double temp,
                 double press_bar, uint64_t seed, double disp_max, double rot_max_deg,
                 double dlnv_max, double p_volume, double *L, double *com, double *quat,
                 double *U, long long *index, int nmoves, long long *counts, int *accepted) {
  g6o_model m;
  g6o_model_get(model, &m);
  const int ns = m.nsite;
  stdterm st[G6O_MAXSITE * G6O_MAXSITE];
  const int nt = std_terms(&m, st);
  npt_cfg g;
  g.k0 = (uint32_t)seed;
  g.k1 = (uint32_t)(seed >> 32);
  g.pvol = p_volume;
  g.dlnv = dlnv_max;
  g.dmax = disp_max / sqrt(3.0);
  g.rot_e = tan(0.5 * rot_max_deg * 3.14159265358979323846 / 180.0) / sqrt(3.0);
  const double beta = 1.0 / temp, press_k = press_bar * 1e5 * 1e-30 / 1.380066e-23;
  double *x = malloc(sizeof(double) * n * ns), *y = malloc(sizeof(double) * n * ns),
         *z = malloc(sizeof(double) * n * ns);
  double *x2 = malloc(sizeof(double) * n * ns), *y2 = malloc(sizeof(double) * n * ns),
         *z2 = malloc(sizeof(double) * n * ns);
  double *com2 = malloc(sizeof(double) * 3 * n), *quat2 = malloc(sizeof(double) * 4 * n);
  double *ep = malloc(sizeof(double) * n * n), *ep2 = malloc(sizeof(double) * n * n);
  int *slot = malloc(sizeof(int) * n), *mem = malloc(sizeof(int) * n);
  double q[4], c[3];
  for (int i = 0; i < n; i++) {
    for (int a = 0; a < 4; a++) q[a] = quat[a * n + i];
    npt_place(&m, q, com + 3 * i, x + i * ns, y + i * ns, z + i * ns);
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      ep[(long)i * n + j] = i == j ? 0.0 : npt_pair(&m, st, nt, com, x, y, z, i, j, *L, rc, taper);
  for (int mv = 0; mv < nmoves; mv++) {
    const unsigned long long idx = (unsigned long long)(*index)++;
    const npt_prop p = npt_make_prop(&g, idx);
    int acc = 0;
    counts[0]++;
    if (p.type && 2.0 * rc >= *L * p.a) {
      /* void: one minimum image per pair needs rc < L'/2 */
      counts[2]++;
      counts[4]++;
    } else if (p.type) {
      counts[2]++;
      const double Lp = *L * p.a;
      for (int k = 0; k < 3 * n; k++) com2[k] = com[k] * p.a;
      for (int i = 0; i < n; i++) {
        for (int a = 0; a < 4; a++) q[a] = quat[a * n + i];
        npt_place(&m, q, com2 + 3 * i, x2 + i * ns, y2 + i * ns, z2 + i * ns);
      }
      double up = 0.0;
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
          const double e = i == j ? 0.0 : npt_pair(&m, st, nt, com2, x2, y2, z2, i, j, Lp, rc, taper);
          ep2[(long)i * n + j] = e;
          if (j > i) up = up + e;
        }
      const double du = up - *U, dv = Lp * Lp * Lp - *L * *L * *L;
      const double arg = -(du * I2J * 0.120273 + press_k * dv) * beta + (double)(n + 1) * p.x;
      acc = arg >= 0.0 || p.uacc < exp(arg);
      if (acc) {
        counts[3]++;
        *L = Lp;
        *U = up;
        memcpy(com, com2, sizeof(double) * 3 * n);
        memcpy(x, x2, sizeof(double) * n * ns);
        memcpy(y, y2, sizeof(double) * n * ns);
        memcpy(z, z2, sizeof(double) * n * ns);
        memcpy(ep, ep2, sizeof(double) * n * n);
      }
    } else {
      int nd = (int)(*L / voxa);
      if (nd < 1) nd = 1;
      const int ncell = nd * nd * nd;
      int cc = (int)(p.a * (double)ncell);
      if (cc > ncell - 1) cc = ncell - 1;
      double off[3];
      npt_offset(&g, idx, *L, nd, off);
      int cnt = 0;
      for (int i = 0; i < n; i++) {
        slot[i] = -1;
        const double sh[3] = {com[3 * i] - off[0], com[3 * i + 1] - off[1], com[3 * i + 2] - off[2]};
        if (npt_cell(sh, *L, nd) == cc) {
          slot[i] = cnt;
          mem[cnt++] = i;
        }
      }
      /* void (rejected, counted): more than 96 members, or a member that the
       * move takes out of the cell, for which the reverse move would pick
       * another member set */
      int out = 0;
      if (cnt <= 96) {
        for (int k = 0; k < cnt; k++) {
          double q1[4], c1[3];
          const int i = mem[k];
          for (int a = 0; a < 4; a++) q[a] = quat[a * n + i];
          npt_member(&g, idx, k, com + 3 * i, q, c1, q1);
          for (int a = 0; a < 3; a++) c1[a] -= off[a];
          if (npt_cell(c1, *L, nd) != cc) out = 1;
        }
      }
      if (cnt > 96 || out) {
        counts[4]++;
      } else {
        memcpy(com2, com, sizeof(double) * 3 * n);
        memcpy(quat2, quat, sizeof(double) * 4 * n);
        memcpy(x2, x, sizeof(double) * n * ns);
        memcpy(y2, y, sizeof(double) * n * ns);
        memcpy(z2, z, sizeof(double) * n * ns);
        for (int k = 0; k < cnt; k++) {
          const int i = mem[k];
          double q1[4];
          for (int a = 0; a < 4; a++) q[a] = quat[a * n + i];
          npt_member(&g, idx, k, com + 3 * i, q, c, q1);
          for (int a = 0; a < 3; a++) com2[3 * i + a] = c[a];
          for (int a = 0; a < 4; a++) quat2[a * n + i] = q1[a];
          npt_place(&m, q1, c, x2 + i * ns, y2 + i * ns, z2 + i * ns);
        }
        double du = 0.0;
        for (int k = 0; k < cnt; k++) {
          const int i = mem[k];
          for (int j = 0; j < n; j++) {
            if (j == i) continue;
            const double e = npt_pair(&m, st, nt, com2, x2, y2, z2, i, j, *L, rc, taper);
            ep2[(long)k * n + j] = e;
            if (slot[j] < 0 || j > i) du = du + (e - ep[(long)i * n + j]);
          }
        }
        const double arg = -(du * I2J * 0.120273) * beta;
        acc = arg >= 0.0 || p.uacc < exp(arg);
        if (acc) {
          *U = *U + du;
          memcpy(com, com2, sizeof(double) * 3 * n);
          memcpy(quat, quat2, sizeof(double) * 4 * n);
          memcpy(x, x2, sizeof(double) * n * ns);
          memcpy(y, y2, sizeof(double) * n * ns);
          memcpy(z, z2, sizeof(double) * n * ns);
          for (int k = 0; k < cnt; k++) {
            const int i = mem[k];
            for (int j = 0; j < n; j++) {
              if (j == i) continue;
              ep[(long)i * n + j] = ep2[(long)k * n + j];
              ep[(long)j * n + i] = ep2[(long)k * n + j];
            }
          }
        }
      }
    }
    counts[1] += acc;
    accepted[mv] = acc;
  }
  free(x); free(y); free(z); free(x2); free(y2); free(z2);
  free(com2); free(quat2); free(ep); free(ep2); free(slot); free(mem);
}

/* ------------------------------------------------------------------------
 * Neighbour arrays (Neighbor.F90).
 * ---------------------------------------------------------------------- */

/* insert (Neighbor.F90:20-88).  ORDER_NONE appends; otherwise the new entry
 * goes before the first entry with a larger key (distance or label), the
 * array growing by one up to `limit` (0: no limit) and losing its last entry
 * when full; an entry whose place is past the limit is dropped. */
static int nb_insert(g6o_nbtab t, int row, int order, int label, double dx, double dy,
                     double dz, double dd, int limit) {
  int *nn = &t.nnear[row];
  const long b = (long)row * G6O_NEAR;
  int pos, last;
  if (order == 0) {
    if (*nn >= G6O_NEAR) return 1;
    pos = (*nn)++;
    t.near[b + pos] = label;
    t.dx[b + pos] = dx;
    t.dy[b + pos] = dy;
    t.dz[b + pos] = dz;
    t.dd[b + pos] = dd;
    return 0;
  }
  for (pos = 0; pos < *nn; pos++) {
    if (order == 1 ? dd < t.dd[b + pos] : label < t.near[b + pos]) break;
  }
  if (limit != 0 && pos >= limit) return 0;
  last = *nn;          /* index the shifted tail ends at */
  if (limit != 0 && last + 1 > limit) last = limit - 1;
  else {
    if (*nn >= G6O_NEAR) return 1;
    (*nn)++;
  }
  for (int j = last; j > pos; j--) {
    t.near[b + j] = t.near[b + j - 1];
    t.dx[b + j] = t.dx[b + j - 1];
    t.dy[b + j] = t.dy[b + j - 1];
    t.dz[b + j] = t.dz[b + j - 1];
    t.dd[b + j] = t.dd[b + j - 1];
  }
  t.near[b + pos] = label;
  t.dx[b + pos] = dx;
  t.dy[b + pos] = dy;
  t.dz[b + pos] = dz;
  t.dd[b + pos] = dd;
  return 0;
}

int g6o_neighbors_all2(int nsite, long npair, const int *pi, const int *pj, const double *ox,
                       const double *oy, const double *oz, const double *coord, int maxnei,
                       int order, double radius, g6o_nbtab out) {
  int err = 0;
  for (int i = 0; i < nsite; i++) out.nnear[i] = 0;
  for (long k = 0; k < npair; k++) {
    const int i = pi[k], j = pj[k];
    double dx = coord[3 * i] - coord[3 * j];
    double dy = coord[3 * i + 1] - coord[3 * j + 1];
    double dz = coord[3 * i + 2] - coord[3 * j + 2];
    dx -= ox[k];
    dy -= oy[k];
    dz -= oz[k];
    const double dd = dx * dx + dy * dy + dz * dz;
    if (radius > 0 && radius * radius < dd) continue;
    err |= nb_insert(out, i, order, j, -dx, -dy, -dz, dd, maxnei);
    err |= nb_insert(out, j, order, i, dx, dy, dz, dd, maxnei);
  }
  return err;
}

int g6o_neighbors_compress(int nrow, g6o_nbtab in, int maxnei, int order, double radius,
                           g6o_nbtab out) {
  int err = 0;
  for (int r = 0; r < nrow; r++) {
    const long b = (long)r * G6O_NEAR;
    out.nnear[r] = 0;
    for (int t = 0; t < in.nnear[r]; t++) {
      if (radius > 0 && radius * radius < in.dd[b + t]) continue;
      err |= nb_insert(out, r, order, in.near[b + t], in.dx[b + t], in.dy[b + t],
                       in.dz[b + t], in.dd[b + t], maxnei);
    }
  }
  return err;
}

/* neighbors_sort (Neighbor.F90:279-333): insertion sort, an entry moving in
 * front of the first entry with a strictly larger key (stable) */
static void nb_sort(g6o_nbtab t, int row, int order) {
  const long b = (long)row * G6O_NEAR;
  const int n = t.nnear[row];
  if (order == 0) return;
  for (int i = 1; i < n; i++) {
    int j;
    for (j = 0; j < i; j++)
      if (order == 1 ? t.dd[b + i] < t.dd[b + j] : t.near[b + i] < t.near[b + j]) break;
    if (j == i) continue;
    const int lab = t.near[b + i];
    const double x = t.dx[b + i], y = t.dy[b + i], z = t.dz[b + i], d = t.dd[b + i];
    for (int k = i; k > j; k--) {
      t.near[b + k] = t.near[b + k - 1];
      t.dx[b + k] = t.dx[b + k - 1];
      t.dy[b + k] = t.dy[b + k - 1];
      t.dz[b + k] = t.dz[b + k - 1];
      t.dd[b + k] = t.dd[b + k - 1];
    }
    t.near[b + j] = lab;
    t.dx[b + j] = x;
    t.dy[b + j] = y;
    t.dz[b + j] = z;
    t.dd[b + j] = d;
  }
}

static void nb_copy_row(g6o_nbtab to, int rt, g6o_nbtab from, int rf) {
  const long bt = (long)rt * G6O_NEAR, bf = (long)rf * G6O_NEAR;
  to.nnear[rt] = from.nnear[rf];
  memcpy(to.near + bt, from.near + bf, sizeof(int) * G6O_NEAR);
  memcpy(to.dx + bt, from.dx + bf, sizeof(double) * G6O_NEAR);
  memcpy(to.dy + bt, from.dy + bf, sizeof(double) * G6O_NEAR);
  memcpy(to.dz + bt, from.dz + bf, sizeof(double) * G6O_NEAR);
  memcpy(to.dd + bt, from.dd + bf, sizeof(double) * G6O_NEAR);
}

int g6o_neighbors_displacement(g6o_nbtab in, int order, int target, double ddx, double ddy,
                               double ddz, int *nmod, int *idx, g6o_nbtab out) {
  /* the moved molecule: every relative vector changes */
  nb_copy_row(out, 0, in, target);
  idx[0] = target;
  for (int t = 0; t < out.nnear[0]; t++) {
    out.dx[t] -= ddx;
    out.dy[t] -= ddy;
    out.dz[t] -= ddz;
    out.dd[t] = out.dx[t] * out.dx[t] + out.dy[t] * out.dy[t] + out.dz[t] * out.dz[t];
  }
  /* its neighbours that list it: the first such entry moves by +delta */
  int m = 1;
  for (int t = 0; t < out.nnear[0]; t++) {
    const int k = out.near[t];
    const long b = (long)k * G6O_NEAR;
    int p;
    for (p = 0; p < in.nnear[k]; p++)
      if (in.near[b + p] == target) break;
    if (p == in.nnear[k]) continue;
    nb_copy_row(out, m, in, k);
    idx[m] = k;
    const long o = (long)m * G6O_NEAR + p;
    out.dx[o] += ddx;
    out.dy[o] += ddy;
    out.dz[o] += ddz;
    out.dd[o] = out.dx[o] * out.dx[o] + out.dy[o] * out.dy[o] + out.dz[o] * out.dz[o];
    m++;
  }
  *nmod = m;
  for (int r = 0; r < m; r++) nb_sort(out, r, order);
  return 0;
}

/* ------------------------------------------------------------------------
 * Umbrella order parameters of mc2 (Q6Local.F90, Triangle.F90, q6.F90,
 * DistribModule.F90), on the neighbour arrays above.
 * ---------------------------------------------------------------------- */

/* plgndr (q6.F90): the associated Legendre function by the upward
 * recurrence in l */
static double q6_plgndr(int l, int m, double x) {
  double pmm = 1.0;
  if (m > 0) {
    const double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double fact = 1.0;
    for (int i = 1; i <= m; i++) {
      pmm = -pmm * fact * somx2;
      fact = fact + 2.0;
    }
  }
  if (l == m) return pmm;
  double pmmp1 = x * (2 * m + 1) * pmm;
  if (l == m + 1) return pmmp1;
  double pll = 0.0;
  for (int ll = m + 2; ll <= l; ll++) {
    pll = (x * (2 * ll - 1) * pmmp1 - (ll + m - 1) * pmm) / (ll - m);
    pmm = pmmp1;
    pmmp1 = pll;
  }
  return pll;
}

/* complex integer power by repeated squaring (the exponent's bits from the
 * lowest), the Fortran runtime's z**m */
static void q6_cpow(double a, double b, int m, double *ra, double *rb) {
  double pa = 1.0, pb = 0.0, xa = a, xb = b;
  int k = m;
  while (k > 0) {
    if (k & 1) {
      const double ta = pa * xa - pb * xb, tb = pa * xb + pb * xa;
      pa = ta;
      pb = tb;
    }
    k >>= 1;
    if (k > 0) {
      const double ta = xa * xa - xb * xb, tb = xa * xb + xb * xa;
      xa = ta;
      xb = tb;
    }
  }
  *ra = pa;
  *rb = pb;
}

/* qlm(6, m, x, y, z) of q6.F90 with ylm2/ylm: w = cmplx(x, y) / |(x, y)|,
 * where cmplx without a kind rounds x and y to single precision and the
 * division is the complex one, (a c + b d, b c - a d) / (c^2 + d^2) with d = 0
 * (checked bit for bit against the reference's qlm built with flang) */
void g6o_q6_qlm(int m, double x, double y, double z, double *re, double *im) {
  static const double prod[13] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0,
                                  362880.0, 3628800.0, 39916800.0, 479001600.0};
  const int l = 6;
  const double r = sqrt(x * x + y * y);
  const double fa = (double)(float)x, fb = (double)(float)y, den = r * r + 0.0 * 0.0;
  const double wa = (fa * r + fb * 0.0) / den, wb = (fb * r - fa * 0.0) / den;
  const int am = m < 0 ? -m : m;
  const double c = sqrt((2.0 * l + 1.0) * prod[l - am] / (4.0 * PI_T * prod[l + am])) *
                   q6_plgndr(l, am, z);
  double pa, pb;
  q6_cpow(wa, wb, am, &pa, &pb);
  double va = c * pa, vb = c * pb;
  if (m < 0) {  /* (-1)**m * conjg(ylm(l, -m)) */
    const double s = (am & 1) ? -1.0 : 1.0;
    va = s * va;
    vb = s * -vb;
  }
  *re = va;
  *im = vb;
}

static int tab_alloc(g6o_nbtab *t, int rows) {
  t->nnear = calloc(rows, sizeof(int));
  t->near = calloc((size_t)rows * G6O_NEAR, sizeof(int));
  t->dx = calloc((size_t)rows * G6O_NEAR * 4, sizeof(double));
  t->dy = t->dx + (size_t)rows * G6O_NEAR;
  t->dz = t->dy + (size_t)rows * G6O_NEAR;
  t->dd = t->dz + (size_t)rows * G6O_NEAR;
  return t->nnear && t->near && t->dx;
}
static void tab_free(g6o_nbtab *t) {
  free(t->nnear);
  free(t->near);
  free(t->dx);
}

/* membership in neis(i) = nei5(i) + i: the label-sorted array or i itself */
static int q6_in(const g6o_nbtab *t5, int i, int k) {
  if (k == i) return 1;
  const int *a = t5->near + (long)i * G6O_NEAR;
  for (int p = 0; p < t5->nnear[i]; p++)
    if (a[p] == k) return 1;
  return 0;
}

/* sq6system_prepareall (Q6Local.F90:101-231): q6[n], nbond[n], the sum and
 * op = sum / n; returns 1 when a molecule has more than MAXBOND = 10 bonds
 * within 3 A (the reference's arrays end there) or an array overflows */
int g6o_q6local(int nsite, long npair, const int *pi, const int *pj, const double *ox,
                const double *oy, const double *oz, const double *coord, double *q6,
                int *nbond, double *sumq6, double *op) {
  g6o_nbtab t6, t3, t5;
  int err = 0;
  if (!tab_alloc(&t6, nsite) || !tab_alloc(&t3, nsite) || !tab_alloc(&t5, nsite)) return 1;
  err |= g6o_neighbors_all2(nsite, npair, pi, pj, ox, oy, oz, coord, 0, 2, 6.0, t6);
  err |= g6o_neighbors_compress(nsite, t6, 0, 2, 3.0, t3);
  err |= g6o_neighbors_compress(nsite, t6, 0, 2, 5.0, t5);
  for (int i = 0; i < nsite; i++) err |= t3.nnear[i] > 10;
  double *qlm = calloc((size_t)nsite * 10 * 26, sizeof(double));
  double *qs = calloc((size_t)nsite * 26, sizeof(double));
  if (!err) {
    for (int i = 0; i < nsite; i++)
      for (int jj = 0; jj < t3.nnear[i]; jj++) {
        const long e = (long)i * G6O_NEAR + jj;
        const double pp = 1.0 / sqrt(t3.dd[e]);
        const double px = t3.dx[e] * pp, py = t3.dy[e] * pp, pz = t3.dz[e] * pp;
        for (int m = -6; m <= 6; m++)
          g6o_q6_qlm(m, px, py, pz, &qlm[((size_t)i * 10 + jj) * 26 + 2 * (m + 6)],
                     &qlm[((size_t)i * 10 + jj) * 26 + 2 * (m + 6) + 1]);
      }
    for (int k = 0; k < nsite; k++) nbond[k] = 0;
    /* neib(jj, i) = neis(i) & neis(j); qlmsum(k) over the bonds whose neib
     * holds k, in the order i, jj */
    for (int i = 0; i < nsite; i++)
      for (int jj = 0; jj < t3.nnear[i]; jj++) {
        const int j = t3.near[(long)i * G6O_NEAR + jj];
        for (int kk = 0; kk <= t5.nnear[i]; kk++) {
          const int k = kk < t5.nnear[i] ? t5.near[(long)i * G6O_NEAR + kk] : i;
          if (!q6_in(&t5, j, k)) continue;
          nbond[k]++;
          for (int c = 0; c < 26; c++)
            qs[(size_t)k * 26 + c] = qs[(size_t)k * 26 + c] + qlm[((size_t)i * 10 + jj) * 26 + c];
        }
      }
    double sum = 0.0;
    for (int i = 0; i < nsite; i++) {
      double nq = 0.0;
      for (int m = 0; m < 13; m++) {
        const double a = hypot(qs[(size_t)i * 26 + 2 * m], qs[(size_t)i * 26 + 2 * m + 1]);
        nq = nq + a * a;
      }
      q6[i] = sqrt(nq * 4.0 * PI_T / (2.0 * 6 + 1.0)) / nbond[i];
      sum = sum + q6[i];
    }
    *sumq6 = sum;
    *op = sum / nsite;
  }
  free(qlm);
  free(qs);
  tab_free(&t6);
  tab_free(&t3);
  tab_free(&t5);
  return err;
}

/* kullback_div (DistribModule.F90:304-326) */
static double kullback_div(int nbin, const double *h1, const double *h2) {
  double s1 = 0.0, s2 = 0.0, div = 0.0;
  for (int i = 0; i < nbin; i++) {
    s1 = s1 + h1[i];
    s2 = s2 + h2[i];
  }
  for (int i = 0; i < nbin; i++) {
    const double p = h1[i] / s1;
    double q = h2[i] / s2;
    if (q == 0) q = p;
    if (p != 0) div = div + p * log(p / q);
  }
  return div;
}

/* sTriangle_prepareall (Triangle.F90:53-116): the O-O-O angles of each
 * molecule's four nearest neighbours within 5 A in a histogram of 21 bins
 * of 0.1 from -1 (histogram_accumulate, DistribModule.F90:149-158), and its
 * Kullback-Leibler divergence from refhist (log of it if logarithmic) */
int g6o_triangle(int nsite, long npair, const int *pi, const int *pj, const double *ox,
                 const double *oy, const double *oz, const double *coord, const double *refhist,
                 int logarithmic, double *hist, double *op) {
  g6o_nbtab t2, t4;
  int err = 0;
  if (!tab_alloc(&t2, nsite) || !tab_alloc(&t4, nsite)) return 1;
  err |= g6o_neighbors_all2(nsite, npair, pi, pj, ox, oy, oz, coord, 0, 1, 5.0, t2);
  err |= g6o_neighbors_compress(nsite, t2, 4, 1, 0.0, t4);
  for (int b = 0; b < 21; b++) hist[b] = 0.0;
  for (int i = 0; i < nsite; i++) {
    const long e = (long)i * G6O_NEAR;
    for (int j = 0; j < t4.nnear[i]; j++)
      for (int k = j + 1; k < t4.nnear[i]; k++) {
        double c = t4.dx[e + j] * t4.dx[e + k] + t4.dy[e + j] * t4.dy[e + k] +
                   t4.dz[e + j] * t4.dz[e + k];
        c = c / sqrt(t4.dd[e + j] * t4.dd[e + k]);
        const int bin = (int)round((c - (-1.0)) / 0.1) + 1;
        if (1 <= bin && bin <= 21) hist[bin - 1] = hist[bin - 1] + 1.0;
      }
  }
  *op = kullback_div(21, hist, refhist);
  if (logarithmic) *op = log(*op);
  tab_free(&t2);
  tab_free(&t4);
  return err;
}

/* ---- Q6 local bond order with its incremental trial update -------------
 * sq6system (Q6Local.F90:42-61) and its trial steps: prepareall (101-231),
 * Differentiate (236-465), Integrate (467-700), Reject (702-750), restated
 * with the same order of every floating-point operation on qlmsum, so that
 * a chain of trials, accepted or rejected, leaves the sums the reference
 * leaves.  Sets (Set.F90) keep their element order: join appends, drop
 * moves the last element into the hole, intersection and difference keep
 * the first set's order. */
#define Q6SET (G6O_NEAR + 1)
#define Q6MAXBOND 10
typedef struct {
  int n;
  g6o_nbtab nei6, nei3, nei5;
  int *neis, *neis_n;       /* [n][Q6SET] */
  int *neib, *neib_n;       /* [n][Q6MAXBOND][Q6SET] */
  double *qlm;              /* [n][Q6MAXBOND][26] */
  double *qlmsum;           /* [n][26] */
  int *nbond;
  double *q6, sumq6;
  /* sQ6diff (Q6Local.F90:47-57) */
  g6o_nbtab new6, new3, new5;
  int nmod, idx[G6O_NEAR + 1];
  int dneis[Q6SET], dneis_n, plus[Q6SET], plus_n, minus[Q6SET], minus_n;
  int dneib[Q6MAXBOND][Q6SET], dneib_n[Q6MAXBOND];
  double dqlm[Q6MAXBOND][26];
  double *dq6, dsumq6;
  int err; /* a trial gave a molecule more than MAXBOND bonds */
} q6st;

static int set_has(const int *a, int na, int x) {
  for (int i = 0; i < na; i++)
    if (a[i] == x) return 1;
  return 0;
}
static void set_join(int *a, int *na, int x) {
  if (!set_has(a, *na, x)) a[(*na)++] = x;
}
static void set_drop(int *a, int *na, int x) {
  for (int i = 0; i < *na; i++)
    if (a[i] == x) {
      a[i] = a[*na - 1];
      (*na)--;
      return;
    }
}
/* intersection (a & b) and difference (a - b), in a's order */
static void set_and(const int *a, int na, const int *b, int nb, int *r, int *nr) {
  *nr = 0;
  for (int i = 0; i < na; i++)
    if (set_has(b, nb, a[i])) r[(*nr)++] = a[i];
}
static void set_minus(const int *a, int na, const int *b, int nb, int *r, int *nr) {
  *nr = 0;
  for (int i = 0; i < na; i++)
    if (!set_has(b, nb, a[i])) r[(*nr)++] = a[i];
}
static int row_lookup(const g6o_nbtab *t, int row, int label) { /* sneighbor1_lookup */
  for (int k = 0; k < t->nnear[row]; k++)
    if (t->near[(long)row * G6O_NEAR + k] == label) return k;
  return -1;
}
#define NEIS(q, i) ((q)->neis + (size_t)(i) * Q6SET)
#define NEIB(q, i, jj) ((q)->neib + ((size_t)(i) * Q6MAXBOND + (jj)) * Q6SET)
#define NEIBN(q, i, jj) ((q)->neib_n[(size_t)(i) * Q6MAXBOND + (jj)])
#define QLM(q, i, jj) ((q)->qlm + ((size_t)(i) * Q6MAXBOND + (jj)) * 26)
#define QSUM(q, k) ((q)->qlmsum + (size_t)(k) * 26)

static void q6_bond_qlm(const g6o_nbtab *t, int row, int jj, double *out) {
  const long e = (long)row * G6O_NEAR + jj;
  const double pp = 1.0 / sqrt(t->dd[e]);
  const double px = t->dx[e] * pp, py = t->dy[e] * pp, pz = t->dz[e] * pp;
  for (int m = -6; m <= 6; m++) g6o_q6_qlm(m, px, py, pz, &out[2 * (m + 6)], &out[2 * (m + 6) + 1]);
}

static double q6_of(const double *qs, int nbond) {
  double nq = 0.0;
  for (int m = 0; m < 13; m++) {
    const double a = hypot(qs[2 * m], qs[2 * m + 1]);
    nq = nq + a * a;
  }
  return sqrt(nq * 4.0 * PI_T / (2.0 * 6 + 1.0)) / nbond;
}

static int q6st_alloc(q6st *q, int n) {
  memset(q, 0, sizeof(*q));
  q->n = n;
  int ok = tab_alloc(&q->nei6, n) && tab_alloc(&q->nei3, n) && tab_alloc(&q->nei5, n) &&
           tab_alloc(&q->new6, G6O_NEAR + 1) && tab_alloc(&q->new3, G6O_NEAR + 1) &&
           tab_alloc(&q->new5, G6O_NEAR + 1);
  q->neis = malloc(sizeof(int) * (size_t)n * Q6SET);
  q->neis_n = calloc(n, sizeof(int));
  q->neib = malloc(sizeof(int) * (size_t)n * Q6MAXBOND * Q6SET);
  q->neib_n = calloc((size_t)n * Q6MAXBOND, sizeof(int));
  q->qlm = calloc((size_t)n * Q6MAXBOND * 26, sizeof(double));
  q->qlmsum = calloc((size_t)n * 26, sizeof(double));
  q->nbond = calloc(n, sizeof(int));
  q->q6 = calloc(n, sizeof(double));
  q->dq6 = calloc(n, sizeof(double));
  return ok && q->neis && q->neis_n && q->neib && q->neib_n && q->qlm && q->qlmsum &&
         q->nbond && q->q6 && q->dq6;
}
static void q6st_free(q6st *q) {
  if (!q->n) return;
  tab_free(&q->nei6); tab_free(&q->nei3); tab_free(&q->nei5);
  tab_free(&q->new6); tab_free(&q->new3); tab_free(&q->new5);
  free(q->neis); free(q->neis_n); free(q->neib); free(q->neib_n); free(q->qlm);
  free(q->qlmsum); free(q->nbond); free(q->q6); free(q->dq6);
}

/* sq6system_prepareall on a pair list; 1 when a molecule has more than
 * MAXBOND bonds or an array overflows */
static int q6_prepare(q6st *q, long np, const int *pi, const int *pj, const double *ox,
                      const double *oy, const double *oz, const double *com, double *op) {
  const int n = q->n;
  int err = 0;
  err |= g6o_neighbors_all2(n, np, pi, pj, ox, oy, oz, com, 0, 2, 6.0, q->nei6);
  err |= g6o_neighbors_compress(n, q->nei6, 0, 2, 3.0, q->nei3);
  err |= g6o_neighbors_compress(n, q->nei6, 0, 2, 5.0, q->nei5);
  if (err) return 1;
  for (int i = 0; i < n; i++) {
    if (q->nei3.nnear[i] > Q6MAXBOND) return 1;
    q->neis_n[i] = 0;
    for (int k = 0; k < q->nei5.nnear[i]; k++) NEIS(q, i)[q->neis_n[i]++] = q->nei5.near[(long)i * G6O_NEAR + k];
    set_join(NEIS(q, i), &q->neis_n[i], i);
  }
  for (int i = 0; i < n; i++)
    for (int jj = 0; jj < q->nei3.nnear[i]; jj++) {
      const int j = q->nei3.near[(long)i * G6O_NEAR + jj];
      set_and(NEIS(q, i), q->neis_n[i], NEIS(q, j), q->neis_n[j], NEIB(q, i, jj), &NEIBN(q, i, jj));
      q6_bond_qlm(&q->nei3, i, jj, QLM(q, i, jj));
    }
  memset(q->qlmsum, 0, sizeof(double) * (size_t)n * 26);
  memset(q->nbond, 0, sizeof(int) * n);
  for (int i = 0; i < n; i++)
    for (int jj = 0; jj < q->nei3.nnear[i]; jj++)
      for (int kk = 0; kk < NEIBN(q, i, jj); kk++) {
        const int k = NEIB(q, i, jj)[kk];
        q->nbond[k] += 1;
        for (int c = 0; c < 26; c++) QSUM(q, k)[c] = QSUM(q, k)[c] + QLM(q, i, jj)[c];
      }
  q->sumq6 = 0.0;
  for (int i = 0; i < n; i++) {
    q->q6[i] = q6_of(QSUM(q, i), q->nbond[i]);
    q->sumq6 = q->sumq6 + q->q6[i];
  }
  *op = q->sumq6 / n;
  return 0;
}

/* qlmsum(:, k) += w * 2 * v (w = +1 or -1), nbond(k) += 2 w */
static void q6_bump(q6st *q, int k, const double *v, int w) {
  q->nbond[k] += 2 * w;
  for (int c = 0; c < 26; c++)
    QSUM(q, k)[c] = w > 0 ? QSUM(q, k)[c] + v[c] * 2 : QSUM(q, k)[c] - v[c] * 2;
}

/* sq6system_Differentiate: changes qlmsum and nbond in place (Integrate
 * keeps, Reject undoes them) and returns the new op */
static double q6_differentiate(q6st *q, int target, double dx, double dy, double dz) {
  const int n = q->n;
  for (int ii = 0; ii < q->nei3.nnear[target]; ii++)
    for (int kk = 0; kk < NEIBN(q, target, ii); kk++)
      q6_bump(q, NEIB(q, target, ii)[kk], QLM(q, target, ii), -1);
  g6o_neighbors_displacement(q->nei6, 2, target, dx, dy, dz, &q->nmod, q->idx, q->new6);
  g6o_neighbors_compress(q->nmod, q->new6, 0, 2, 3.0, q->new3);
  g6o_neighbors_compress(q->nmod, q->new6, 0, 2, 5.0, q->new5);
  for (int r = 0; r < q->nmod; r++) q->err |= q->new3.nnear[r] > Q6MAXBOND;
  q->dneis_n = 0;
  for (int k = 0; k < q->new5.nnear[0]; k++) q->dneis[q->dneis_n++] = q->new5.near[k];
  set_join(q->dneis, &q->dneis_n, target);
  set_minus(q->dneis, q->dneis_n, NEIS(q, target), q->neis_n[target], q->plus, &q->plus_n);
  set_minus(NEIS(q, target), q->neis_n[target], q->dneis, q->dneis_n, q->minus, &q->minus_n);
  for (int jj = 0; jj < q->new3.nnear[0] && jj < Q6MAXBOND; jj++) {
    const int j = q->new3.near[jj];
    set_and(q->dneis, q->dneis_n, NEIS(q, j), q->neis_n[j], q->dneib[jj], &q->dneib_n[jj]);
    q6_bond_qlm(&q->new3, 0, jj, q->dqlm[jj]);
  }
  for (int jj = 0; jj < q->new3.nnear[0] && jj < Q6MAXBOND; jj++)
    for (int kk = 0; kk < q->dneib_n[jj]; kk++) q6_bump(q, q->dneib[jj][kk], q->dqlm[jj], 1);
  for (int ii = 0; ii < q->plus_n; ii++) {
    const int i = q->plus[ii];
    for (int jj = 0; jj < q->nei3.nnear[i]; jj++)
      if (set_has(q->dneis, q->dneis_n, q->nei3.near[(long)i * G6O_NEAR + jj]))
        q6_bump(q, target, QLM(q, i, jj), 1);
  }
  for (int ii = 0; ii < q->minus_n; ii++) {
    const int i = q->minus[ii];
    for (int jj = 0; jj < q->nei3.nnear[i]; jj++)
      if (set_has(NEIS(q, target), q->neis_n[target], q->nei3.near[(long)i * G6O_NEAR + jj]))
        q6_bump(q, target, QLM(q, i, jj), -1);
  }
  q->dsumq6 = 0.0;
  for (int i = 0; i < n; i++) {
    q->dq6[i] = q6_of(QSUM(q, i), q->nbond[i]);
    q->dsumq6 = q->dsumq6 + q->dq6[i];
  }
  return q->dsumq6 / n;
}

static void copy_row(g6o_nbtab *dst, int drow, const g6o_nbtab *src, int srow);

/* sq6system_Integrate */
static void q6_integrate(q6st *q, int target) {
  const int n = q->n;
  memcpy(q->q6, q->dq6, sizeof(double) * n);
  q->sumq6 = q->dsumq6;
  for (int jj = 0; jj < q->new3.nnear[0] && jj < Q6MAXBOND; jj++)
    memcpy(QLM(q, target, jj), q->dqlm[jj], sizeof(double) * 26);
  memcpy(NEIS(q, target), q->dneis, sizeof(int) * q->dneis_n);
  q->neis_n[target] = q->dneis_n;
  for (int ii = 0; ii < q->plus_n; ii++) set_join(NEIS(q, q->plus[ii]), &q->neis_n[q->plus[ii]], target);
  for (int ii = 0; ii < q->minus_n; ii++) set_drop(NEIS(q, q->minus[ii]), &q->neis_n[q->minus[ii]], target);
  int oldb[Q6SET], newb[Q6SET], lost[Q6SET], found[Q6SET], kept[Q6SET];
  int no = q->nei3.nnear[target], nn = q->new3.nnear[0], nl, nf, nk;
  memcpy(oldb, q->nei3.near + (long)target * G6O_NEAR, sizeof(int) * no);
  memcpy(newb, q->new3.near, sizeof(int) * nn);
  set_minus(oldb, no, newb, nn, lost, &nl);
  set_minus(newb, nn, oldb, no, found, &nf);
  set_and(oldb, no, newb, nn, kept, &nk);
  for (int ii = 0; ii < q->plus_n; ii++) {
    const int i = q->plus[ii];
    for (int jj = 0; jj < q->nei3.nnear[i]; jj++) {
      const int j = q->nei3.near[(long)i * G6O_NEAR + jj];
      if (!set_has(q->dneis, q->dneis_n, j)) continue;
      set_join(NEIB(q, i, jj), &NEIBN(q, i, jj), target);
      const int kk = row_lookup(&q->nei3, j, i);
      if (kk >= 0) set_join(NEIB(q, j, kk), &NEIBN(q, j, kk), target);
    }
  }
  for (int ii = 0; ii < q->minus_n; ii++) {
    const int i = q->minus[ii];
    for (int jj = 0; jj < q->nei3.nnear[i]; jj++) {
      const int j = q->nei3.near[(long)i * G6O_NEAR + jj];
      /* neis(target) is the new set here (assigned above) */
      if (!set_has(NEIS(q, target), q->neis_n[target], j)) continue;
      set_drop(NEIB(q, i, jj), &NEIBN(q, i, jj), target);
      const int kk = row_lookup(&q->nei3, j, i);
      if (kk >= 0) set_drop(NEIB(q, j, kk), &NEIBN(q, j, kk), target);
    }
  }
  for (int r = 0; r < q->nmod; r++) copy_row(&q->nei6, q->idx[r], &q->new6, r);
  for (int ii = 0; ii < nl; ii++) { /* lost bonds: the target's qlm drops out */
    const int i = lost[ii];
    int jj = 0;
    for (int kk = 0; kk < q->nei3.nnear[i]; kk++)
      if (q->nei3.near[(long)i * G6O_NEAR + kk] != target) {
        memmove(QLM(q, i, jj), QLM(q, i, kk), sizeof(double) * 26);
        jj++;
      }
  }
  double tmp2[26] = {0};
  for (int ii = 0; ii < nk; ii++) { /* kept bonds: the target's new qlm */
    const int i = kept[ii];
    const int jj = row_lookup(&q->new3, 0, i);
    if (jj >= 0) memcpy(tmp2, q->dqlm[jj], sizeof(tmp2));
    const int kk = row_lookup(&q->nei3, i, target);
    if (kk >= 0) memcpy(QLM(q, i, kk), tmp2, sizeof(tmp2));
  }
  for (int ii = 0; ii < q->nmod; ii++) { /* found bonds: room for the target's */
    const int i = q->idx[ii];
    if (!set_has(found, nf, i)) continue;
    double tmp[Q6MAXBOND][26];
    const int m = q->new3.nnear[ii];
    for (int c = 0; c + 1 < m && c < Q6MAXBOND; c++) memcpy(tmp[c], QLM(q, i, c), sizeof(tmp[c]));
    const int jj0 = row_lookup(&q->new3, 0, i);
    if (jj0 >= 0) memcpy(tmp2, q->dqlm[jj0], sizeof(tmp2));
    int jj = 0;
    for (int kk = 0; kk < m && kk < Q6MAXBOND; kk++) {
      if (q->new3.near[(long)ii * G6O_NEAR + kk] != target) {
        memcpy(QLM(q, i, kk), tmp[jj], sizeof(tmp2));
        jj++;
      } else {
        memcpy(QLM(q, i, kk), tmp2, sizeof(tmp2));
      }
    }
  }
  for (int r = 0; r < q->nmod; r++) {
    copy_row(&q->nei3, q->idx[r], &q->new3, r);
    copy_row(&q->nei5, q->idx[r], &q->new5, r);
  }
  for (int jj = 0; jj < q->nei3.nnear[target]; jj++) {
    const int j = q->nei3.near[(long)target * G6O_NEAR + jj];
    set_and(NEIS(q, target), q->neis_n[target], NEIS(q, j), q->neis_n[j], NEIB(q, target, jj),
            &NEIBN(q, target, jj));
    for (int kk = 0; kk < q->nei3.nnear[j]; kk++) {
      const int k = q->nei3.near[(long)j * G6O_NEAR + kk];
      set_and(NEIS(q, j), q->neis_n[j], NEIS(q, k), q->neis_n[k], NEIB(q, j, kk), &NEIBN(q, j, kk));
    }
  }
  for (int ii = 0; ii < nl; ii++) {
    const int i = lost[ii];
    for (int jj = 0; jj < q->nei3.nnear[i]; jj++) {
      const int j = q->nei3.near[(long)i * G6O_NEAR + jj];
      set_and(NEIS(q, i), q->neis_n[i], NEIS(q, j), q->neis_n[j], NEIB(q, i, jj), &NEIBN(q, i, jj));
    }
  }
}

/* sq6system_Reject: the sums of Differentiate taken back, in its order */
static void q6_reject(q6st *q, int target) {
  for (int jj = 0; jj < q->new3.nnear[0] && jj < Q6MAXBOND; jj++)
    for (int kk = 0; kk < q->dneib_n[jj]; kk++) q6_bump(q, q->dneib[jj][kk], q->dqlm[jj], -1);
  for (int ii = 0; ii < q->nei3.nnear[target]; ii++)
    for (int kk = 0; kk < NEIBN(q, target, ii); kk++)
      q6_bump(q, NEIB(q, target, ii)[kk], QLM(q, target, ii), 1);
  for (int ii = 0; ii < q->plus_n; ii++) {
    const int i = q->plus[ii];
    for (int jj = 0; jj < q->nei3.nnear[i]; jj++)
      if (set_has(q->dneis, q->dneis_n, q->nei3.near[(long)i * G6O_NEAR + jj]))
        q6_bump(q, target, QLM(q, i, jj), -1);
  }
  for (int ii = 0; ii < q->minus_n; ii++) {
    const int i = q->minus[ii];
    for (int jj = 0; jj < q->nei3.nnear[i]; jj++)
      if (set_has(NEIS(q, target), q->neis_n[target], q->nei3.near[(long)i * G6O_NEAR + jj]))
        q6_bump(q, target, QLM(q, i, jj), 1);
  }
}

/* ---- mc2's trial loop with umbrellas ------------------------------------
 * Umbrella potentials coeffi (op - center)^2 (Umbrella.F90:11-16) on the
 * order parameters of @GEUM (mc2util.F90:174-191): the triangle histogram's
 * divergence (TRIANGLE, LOGTRIAN) and the energy (EPOP____).  Each call is
 * one loop of mc2 (mc2.F90:407-899): the pair list of the loop (all pairs
 * within rc, mc2.F90:486-496), prepareall (mc2.F90:509-521), then every
 * trial adds sum(U(newop) - U(op)) to deltae (mc2.F90:684-706) and an
 * accepted trial integrates the tables (mc2.F90:754-771). */
struct umbctx {
  int numb;
  const int *type;
  const double *center, *coeffi, *refhist;
  double *op;
  double newop[G6O_MAXUMB];
  /* sTriangle per umbrella: nei2 (5 A shell by distance), nei (nearest 4),
   * hist; the trial's modified arrays and histogram */
  g6o_nbtab nei2[G6O_MAXUMB], nei[G6O_MAXUMB], new2, new4;
  double hist[G6O_MAXUMB][21], newhist[G6O_MAXUMB][21];
  int nmod, idx[G6O_NEAR + 1], n;
  char *inset;
  q6st q6[G6O_MAXUMB]; /* sq6system per Q6LOCAL_ umbrella */
  int target;
};

static void hist_add(double *h, const g6o_nbtab *t, int row, double amount) {
  const long e = (long)row * G6O_NEAR;
  for (int j = 0; j < t->nnear[row]; j++)
    for (int k = j + 1; k < t->nnear[row]; k++) {
      double c = t->dx[e + j] * t->dx[e + k] + t->dy[e + j] * t->dy[e + k] +
                 t->dz[e + j] * t->dz[e + k];
      c = c / sqrt(t->dd[e + j] * t->dd[e + k]);
      const int bin = (int)round((c - (-1.0)) / 0.1) + 1;
      if (1 <= bin && bin <= 21) h[bin - 1] = h[bin - 1] + amount;
    }
}

static int umb_tri(const umbctx *U, int b) { return U->type[b] == 1 || U->type[b] == 4; }

/* prepareall: sTriangle_prepareall (Triangle.F90:53-116) on the loop's pair
 * list; op = epsum for EPOP____ (mc2.F90:519-520) */
static int umb_prepare(umbctx *U, int n, double L, double rc, const double *com, double epsum) {
  int err = 0;
  long np = 0;
  int *pi = malloc(sizeof(int) * (size_t)n * (n - 1) / 2 + 4),
      *pj = malloc(sizeof(int) * (size_t)n * (n - 1) / 2 + 4);
  double *o = malloc(sizeof(double) * 3 * ((size_t)n * (n - 1) / 2 + 1));
  /* the all-pairs list compressed to rc (Interaction_Compress), i < j */
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      double d[3], s[3];
      for (int a = 0; a < 3; a++) {
        d[a] = com[3 * i + a] - com[3 * j + a];
        s[a] = round(d[a] * (1.0 / L)) * L;
        d[a] = d[a] - s[a];
      }
      if (!(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < rc * rc)) continue;
      pi[np] = i;
      pj[np] = j;
      o[3 * np] = s[0];
      o[3 * np + 1] = s[1];
      o[3 * np + 2] = s[2];
      np++;
    }
  double *ox = malloc(sizeof(double) * (np + 1)), *oy = malloc(sizeof(double) * (np + 1)),
         *oz = malloc(sizeof(double) * (np + 1));
  for (long k = 0; k < np; k++) {
    ox[k] = o[3 * k];
    oy[k] = o[3 * k + 1];
    oz[k] = o[3 * k + 2];
  }
  U->n = n;
  for (int b = 0; b < U->numb; b++) {
    if (U->type[b] == 3) {
      U->op[b] = epsum;
    } else if (U->type[b] == 2) {
      err |= q6_prepare(&U->q6[b], np, pi, pj, ox, oy, oz, com, &U->op[b]);
    } else if (umb_tri(U, b)) {
      err |= g6o_neighbors_all2(n, np, pi, pj, ox, oy, oz, com, 0, 1, 5.0, U->nei2[b]);
      err |= g6o_neighbors_compress(n, U->nei2[b], 4, 1, 0.0, U->nei[b]);
      for (int k = 0; k < 21; k++) U->hist[b][k] = 0.0;
      for (int i = 0; i < n; i++) hist_add(U->hist[b], &U->nei[b], i, 1.0);
      U->op[b] = kullback_div(21, U->hist[b], U->refhist);
      if (U->type[b] == 4) U->op[b] = log(U->op[b]);
    } else {
      err = 1;
    }
  }
  free(pi); free(pj); free(o); free(ox); free(oy); free(oz);
  return err;
}

static int row_has(const g6o_nbtab *t, int row, int label) {
  for (int k = 0; k < t->nnear[row]; k++)
    if (t->near[(long)row * G6O_NEAR + k] == label) return 1;
  return 0;
}

/* sTriangle_Differentiate (Triangle.F90:125-249): the modified arrays, the
 * set of molecules whose angles may change, the histogram with the set's
 * old angles taken out (partial_accumulate) and the new ones put in
 * (partial_accumulate_indexed, DistribModule.F90:231-300) */
static double tri_differentiate(umbctx *U, int b, int target, double dx, double dy, double dz) {
  g6o_nbtab *n2 = &U->nei2[b], *n4 = &U->nei[b];
  g6o_neighbors_displacement(*n2, 1, target, dx, dy, dz, &U->nmod, U->idx, U->new2);
  g6o_neighbors_compress(U->nmod, U->new2, 4, 1, 0.0, U->new4);
  int list[2 * G6O_NEAR + 8], ns = 0;
#define JOIN(v)                \
  do {                         \
    const int v_ = (v);        \
    if (!U->inset[v_]) {       \
      U->inset[v_] = 1;        \
      list[ns++] = v_;         \
    }                          \
  } while (0)
  JOIN(target);
  for (int k = 0; k < n4->nnear[target]; k++) JOIN(n4->near[(long)target * G6O_NEAR + k]);
  for (int k = 0; k < n2->nnear[target]; k++) {
    const int i = n2->near[(long)target * G6O_NEAR + k];
    if (row_has(n4, i, target)) JOIN(i);
  }
  for (int k = 0; k < U->new4.nnear[0]; k++) JOIN(U->new4.near[k]);
  for (int r = 1; r < U->nmod; r++)
    if (row_has(&U->new4, r, target)) JOIN(U->idx[r]);
#undef JOIN
  memcpy(U->newhist[b], U->hist[b], sizeof(double) * 21);
  for (int k = 0; k < ns; k++) hist_add(U->newhist[b], n4, list[k], -1.0);
  for (int r = 0; r < U->nmod; r++)
    if (U->inset[U->idx[r]]) hist_add(U->newhist[b], &U->new4, r, 1.0);
  for (int k = 0; k < ns; k++) U->inset[list[k]] = 0;
  double op = kullback_div(21, U->newhist[b], U->refhist);
  if (U->type[b] == 4) op = log(op);
  return op;
}

static void copy_row(g6o_nbtab *dst, int drow, const g6o_nbtab *src, int srow) {
  const long d = (long)drow * G6O_NEAR, s = (long)srow * G6O_NEAR;
  const int m = src->nnear[srow];
  dst->nnear[drow] = m;
  memcpy(dst->near + d, src->near + s, sizeof(int) * m);
  memcpy(dst->dx + d, src->dx + s, sizeof(double) * m);
  memcpy(dst->dy + d, src->dy + s, sizeof(double) * m);
  memcpy(dst->dz + d, src->dz + s, sizeof(double) * m);
  memcpy(dst->dd + d, src->dd + s, sizeof(double) * m);
}

static double umbrella_potential(double coeffi, double center, double op) {
  return coeffi * ((op - center) * (op - center));
}

/* every umbrella's newop, then deltaesum = deltae + deltaope
 * (mc2.F90:684-706); the triangle umbrellas keep their modified arrays
 * until the decision.  Umbrellas of one kind on one configuration evolve
 * alike, but each keeps its own tables as the reference's sTriangle does;
 * the arrays of the trial are those of the last triangle umbrella, which
 * are the same for all of them. */
static double umb_trial(umbctx *U, int target, double dx, double dy, double dz, double epsum,
                        double deltae) {
  U->target = target;
  for (int b = 0; b < U->numb; b++)
    U->newop[b] = U->type[b] == 3   ? epsum + deltae
                  : U->type[b] == 2 ? q6_differentiate(&U->q6[b], target, dx, dy, dz)
                                    : tri_differentiate(U, b, target, dx, dy, dz);
  double deltaope = 0.0;
  for (int b = 0; b < U->numb; b++)
    deltaope = deltaope + umbrella_potential(U->coeffi[b], U->center[b], U->newop[b]) -
               umbrella_potential(U->coeffi[b], U->center[b], U->op[b]);
  return deltae + deltaope;
}

/* sTriangle_Integrate (Triangle.F90:251-280) and op = newop */
static void umb_accept(umbctx *U) {
  for (int b = 0; b < U->numb; b++) {
    if (umb_tri(U, b)) {
      for (int r = 0; r < U->nmod; r++) {
        copy_row(&U->nei2[b], U->idx[r], &U->new2, r);
        copy_row(&U->nei[b], U->idx[r], &U->new4, r);
      }
      memcpy(U->hist[b], U->newhist[b], sizeof(double) * 21);
    } else if (U->type[b] == 2) {
      q6_integrate(&U->q6[b], U->target);
    }
    U->op[b] = U->newop[b];
  }
}

/* the rejected trial: sq6system_Reject (mc2.F90:820-829) */
static void umb_reject(umbctx *U) {
  for (int b = 0; b < U->numb; b++)
    if (U->type[b] == 2) q6_reject(&U->q6[b], U->target);
}

static int umb_failed(const umbctx *U) {
  for (int b = 0; b < U->numb; b++)
    if (U->type[b] == 2 && U->q6[b].err) return 1;
  return 0;
}

long g6o_mc_trials_umbrella(int model, int n, double L, double rc, double taper, double temp,
                            double *com, double *quat, double *delta, double *rtr, double *ep,
                            double *epsum, const double *u, long *pos, int ntrial, int numb,
                            const int *type, const double *center, const double *coeffi,
                            const double *refhist, double *op) {
  if (numb < 0 || numb > G6O_MAXUMB) return -1;
  umbctx U;
  memset(&U, 0, sizeof(U));
  U.numb = numb;
  U.type = type;
  U.center = center;
  U.coeffi = coeffi;
  U.refhist = refhist;
  U.op = op;
  int ok = tab_alloc(&U.new2, G6O_NEAR + 1) && tab_alloc(&U.new4, G6O_NEAR + 1);
  for (int b = 0; b < numb; b++) {
    if (umb_tri(&U, b)) ok = ok && tab_alloc(&U.nei2[b], n) && tab_alloc(&U.nei[b], n);
    if (type[b] == 2) ok = ok && q6st_alloc(&U.q6[b], n);
  }
  U.inset = calloc(n, 1);
  long acc = -1;
  if (ok && U.inset)
    acc = mc_trials(model, n, L, rc, taper, temp, com, quat, delta, rtr, ep, epsum, u, pos,
                    ntrial, &U);
  if (umb_failed(&U)) acc = -1;
  for (int b = 0; b < numb; b++) {
    if (U.nei2[b].nnear) {
      tab_free(&U.nei2[b]);
      tab_free(&U.nei[b]);
    }
    q6st_free(&U.q6[b]);
  }
  tab_free(&U.new2);
  tab_free(&U.new4);
  free(U.inset);
  return acc;
}
