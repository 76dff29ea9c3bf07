// C ABI of g6gpu (include/g6gpu.h): handles, host-side tables, state
// transfer, the MD step driver (stream launches captured into a hipGraph),
// MT19937 for the Maxwell-Boltzmann start, and the synthetic lattice start.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "g6_internal.h"

namespace g6 {

// ---- reference constants (PhysConst.F90, TIP4PConst.F90, ST2Module.F90) ---
static const double kPI = 3.1415926535897932;  // PItanaka
static const double kNA = 6.0225e23, kKB = 1.380066e-23, kCA = 4.184;
static double coeff() { return kCA * 332.17752; }
static double ipow(double x, int k) {
  double r = x;
  for (int i = 1; i < k; i++) r = r * x;
  return r;
}

// the reciprocals rigid_correct multiplies by (1/m, 1/I), rounded once
static void set_inverses(ModelTable *m) {
  m->inv_mass = 1.0 / m->mass;
  m->inv_ixx = 1.0 / m->Ixx;
  m->inv_iyy = 1.0 / m->Iyy;
  m->inv_izz = 1.0 / m->Izz;
}

// StdForce's pair coefficients (StdInteraction.F90:206-221) from the site
// parameters std_q, std_eps, std_sig for the Coulomb site pairs ci/cj and
// the O-O LJ term, in place of a custom kernel's constants
static void std_coefficients(ModelTable *m) {
  const double KX2X = 1.0 / 1e-3, J2I = 1.0 / 10.0;
  for (int t = 0; t < m->ncoul; t++)
    m->qq[t] = m->std_q[m->ci[t]] * m->std_q[m->cj[t]] * coeff() * KX2X * J2I * 1.0;
  const double eps = m->std_eps[0], sig = m->std_sig[0];
  const double ljeps = sqrt(eps * eps);
  const double s6 = ipow((sig + sig) * 0.5, 6);
  m->lj_a = 4.0 * ljeps * s6 * s6 * KX2X * J2I * 1.0;
  m->lj_b = -4.0 * ljeps * s6 * KX2X * J2I * 1.0;
  m->rl = m->ru = 0.0;
}

bool make_model_table(int model, int prec, ModelTable *m, bool fast) {
  (void)prec;
  memset(m, 0, sizeof(*m));
  m->model = model;
  m->lj_site = 0;
  if (model == G6_TIP4P) {
    // Rigid_TIP4P_Constructor (TIP4PModule.F90:70-97)
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
    // custom-kernel constants WWMM, WWMH, WWHH, AA, BB (TIP4PConst.F90:4-13)
    const double C = coeff();
    const double wwmm = -1.04 * -1.04 * C * 100.0, wwmh = -1.04 * 0.52 * C * 100.0,
                 wwhh = 0.52 * 0.52 * C * 100.0;
    const int ci[9] = {1, 1, 1, 2, 3, 2, 2, 3, 3}, cj[9] = {1, 2, 3, 1, 1, 2, 3, 2, 3};
    m->ncoul = 9;
    for (int t = 0; t < 9; t++) {
      m->ci[t] = ci[t];
      m->cj[t] = cj[t];
      m->qq[t] = (ci[t] == 1 && cj[t] == 1) ? wwmm
                 : (ci[t] == 1 || cj[t] == 1) ? wwmh : wwhh;
    }
    m->lj_a = 600000.0 * kCA * 100.0;
    m->lj_b = -610.0 * kCA * 100.0;
    // tip4p_setinteraction (TIP4PModule.F90:43-68): charges WM, WH; O LJ
    // EPSWO, SIGWO (TIP4PConst.F90:7-17)
    m->std_q[1] = -1.04;
    m->std_q[2] = m->std_q[3] = 0.52;
    m->std_eps[0] = 0.15504166666666666666666 * kCA;
    m->std_sig[0] = 3.15357794197649532464;
    if (!fast) std_coefficients(m);  // StdForce (StdInteraction.F90:206-221)
    set_inverses(m);
    return true;
  }
  if (model == G6_TIP5P || model == G6_ST2) {
    const bool st2 = model == G6_ST2;
    // Rigid_TIP5P_Constructor (RigidModule.F90:429-477),
    // Rigid_ST2_Constructor (ST2Module.F90:60-100)
    const double theta = (st2 ? 109.47 : 104.52) * kPI / 180.0;
    const double phi = 109.47 * kPI / 180.0;
    const double l1 = st2 ? 1.0 : 0.9572, l2 = st2 ? 0.8 : 0.70;
    const double hmass = 1.0, omass = 16.0;
    const double ang2 = st2 ? theta : phi;
    m->nsite = 6;
    m->by[1] = l1 * sin(0.5 * theta);
    m->bz[1] = l1 * cos(0.5 * theta);
    m->by[2] = -l1 * sin(0.5 * theta);
    m->bz[2] = l1 * cos(0.5 * theta);
    m->bx[3] = l2 * sin(0.5 * ang2);
    m->bz[3] = -l2 * cos(0.5 * ang2);
    m->bx[4] = -l2 * sin(0.5 * ang2);
    m->bz[4] = -l2 * cos(0.5 * ang2);
    const double zshift = m->bz[1] * 2.0 * hmass / (2.0 * hmass + omass);
    for (int i = 0; i < 5; i++) m->bz[i] = m->bz[i] - zshift;
    m->Ixx = omass * (m->by[0] * m->by[0] + m->bz[0] * m->bz[0]) +
             hmass * (m->by[1] * m->by[1] + m->bz[1] * m->bz[1]) * 2.0;
    m->Iyy = omass * (m->bz[0] * m->bz[0] + m->bx[0] * m->bx[0]) +
             hmass * (m->bz[1] * m->bz[1] + m->bx[1] * m->bx[1]) * 2.0;
    m->Izz = omass * (m->bx[0] * m->bx[0] + m->by[0] * m->by[0]) +
             hmass * (m->bx[1] * m->bx[1] + m->by[1] * m->by[1]) * 2.0;
    m->mass = 2.0 * hmass + omass;
    m->ncoul = 16;
    if (!st2) {
      // tip5p_setinteraction (TIP5PModule.F90:57-83)
      m->std_q[1] = m->std_q[2] = 0.2410;
      m->std_q[3] = m->std_q[4] = -0.2410;
      m->std_eps[0] = 0.6694;
      m->std_sig[0] = 3.12;
      // StdForce coefficients (StdInteraction.F90:206-221), the live TIP5P path
      const double q[6] = {0.0, 0.2410, 0.2410, -0.2410, -0.2410, 0.0};
      const double KX2X = 1.0 / 1e-3, J2I = 1.0 / 10.0;
      for (int t = 0; t < 16; t++) {
        const int a = 1 + t / 4, b = 1 + t % 4;
        m->ci[t] = a;
        m->cj[t] = b;
        m->qq[t] = q[a] * q[b] * coeff() * KX2X * J2I * 1.0;
      }
      const double eps = 0.6694, sig = 3.12;
      const double ljeps = sqrt(eps * eps);
      const double s6 = ipow((sig + sig) * 0.5, 6);
      m->lj_a = 4.0 * ljeps * s6 * s6 * KX2X * J2I * 1.0;
      m->lj_b = -4.0 * ljeps * s6 * KX2X * J2I * 1.0;
    } else {
      // force_st2 constants (ST2Module.F90:15-22, 121-125); the charges and
      // eps are single-precision literals in the reference
      const double qh = (double)0.2357f, qm = (double)(-0.2357f);
      const double eps = (double)0.52605e-21f * kNA * 1e-3, sig = 3.100;
      const double C = coeff();
      for (int t = 0; t < 16; t++) {
        const int a = 1 + t / 4, b = 1 + t % 4;
        m->ci[t] = a;
        m->cj[t] = b;
        const bool ha = a <= 2, hb = b <= 2;
        m->qq[t] = (ha && hb) ? qh * qh * C * 100.0
                   : (!ha && !hb) ? qm * qm * C * 100.0 : qh * qm * C * 100.0;
      }
      m->lj_a = 4.0 * eps * ipow(sig, 12) * 100.0;
      m->lj_b = -4.0 * eps * ipow(sig, 6) * 100.0;
      m->rl = 2.0160;
      m->ru = 3.1287;
      // st2_setinteraction (ST2Module.F90:32-58): the StdForce parameters
      m->std_q[1] = m->std_q[2] = qh;
      m->std_q[3] = m->std_q[4] = qm;
      m->std_eps[0] = eps;
      m->std_sig[0] = sig;
      if (!fast) std_coefficients(m);  // StdForce: no Stillinger-Rahman switch
    }
    set_inverses(m);
    return true;
  }
  if (model == G6_SPCE) {
    // Rigid_SPCE_Constructor (SPCEModule.F90:48-80): O, H, H, CoM
    const double theta = 109.5 * kPI / 180.0, l1 = 1.0, hmass = 1.0, omass = 16.0;
    m->nsite = 4;
    m->by[1] = l1 * sin(0.5 * theta);
    m->bz[1] = l1 * cos(0.5 * theta);
    m->by[2] = -l1 * sin(0.5 * theta);
    m->bz[2] = l1 * cos(0.5 * theta);
    const double zshift = m->bz[1] * 2.0 * hmass / (2.0 * hmass + omass);
    for (int i = 0; i < 3; i++) m->bz[i] = m->bz[i] - zshift;
    m->Ixx = omass * (m->by[0] * m->by[0] + m->bz[0] * m->bz[0]) +
             hmass * (m->by[1] * m->by[1] + m->bz[1] * m->bz[1]) * 2.0;
    m->Iyy = omass * (m->bz[0] * m->bz[0] + m->bx[0] * m->bx[0]) +
             hmass * (m->bz[1] * m->bz[1] + m->bx[1] * m->bx[1]) * 2.0;
    m->Izz = omass * (m->bx[0] * m->bx[0] + m->by[0] * m->by[0]) +
             hmass * (m->bx[1] * m->bx[1] + m->by[1] * m->by[1]) * 2.0;
    m->mass = 2.0 * hmass + omass;
    // spce_setinteraction (SPCEModule.F90:25-46): the O site carries the
    // charge and the LJ term; eps from K by J2K (PhysConst.F90:56)
    const double q[3] = {-0.8476, 0.4238, 0.4238};
    const double eps = 78.208 / 0.120273 * 1e-3, sig = 3.166;
    m->std_q[0] = q[0];
    m->std_q[1] = m->std_q[2] = q[1];
    m->std_eps[0] = eps;
    m->std_sig[0] = sig;
    // StdForce coefficients (StdInteraction.F90:206-221): 9 Coulomb terms
    // over O, H, H and the O-O LJ term
    const double KX2X = 1.0 / 1e-3, J2I = 1.0 / 10.0;
    m->ncoul = 9;
    for (int t = 0; t < 9; t++) {
      const int a = t / 3, b = t % 3;
      m->ci[t] = a;
      m->cj[t] = b;
      m->qq[t] = q[a] * q[b] * coeff() * KX2X * J2I * 1.0;
    }
    const double ljeps = sqrt(eps * eps);
    const double s6 = ipow((sig + sig) * 0.5, 6);
    m->lj_a = 4.0 * ljeps * s6 * s6 * KX2X * J2I * 1.0;
    m->lj_b = -4.0 * ljeps * s6 * KX2X * J2I * 1.0;
    set_inverses(m);
    return true;
  }
  if (model == G6_NVDE) {
    // Rigid_NvdE_Constructor (NvdEModule.F90:90-148): O, H, H, M, L, L, CoM;
    // the x shift is comz/mass as the constructor computes it
    const double oh = 0.980, om = 0.230, ol = 0.8892;
    const double hoh = 108.00 * kPI / 180.0, lol = 111.0 * kPI / 180.0;
    const double ms[7] = {16.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0};
    double x[7] = {0, 0, 0, 0, ol * sin(0.5 * lol), -ol * sin(0.5 * lol), 0};
    double y[7] = {0, oh * sin(0.5 * hoh), -oh * sin(0.5 * hoh), 0, 0, 0, 0};
    double z[7] = {0, oh * cos(0.5 * hoh), oh * cos(0.5 * hoh), om, -ol * cos(0.5 * lol),
                   -ol * cos(0.5 * lol), 0};
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
    m->Ixx = m->Iyy = m->Izz = 0.0;
    for (int i = 0; i < 7; i++) {
      m->bx[i] = x[i];
      m->by[i] = y[i];
      m->bz[i] = z[i];
      m->Ixx = m->Ixx + ms[i] * (y[i] * y[i] + z[i] * z[i]);
      m->Iyy = m->Iyy + ms[i] * (z[i] * z[i] + x[i] * x[i]);
      m->Izz = m->Izz + ms[i] * (x[i] * x[i] + y[i] * y[i]);
    }
    m->mass = mass;
    // nvde_setinteraction (NvdEModule.F90:58-88), NvdEModule.F90:10-21
    const double rgas = kNA * kKB;
    const double q[6] = {0.0, 0.477, 0.477, -0.866, -0.044, -0.044};
    const double eps[3] = {85.9766 * rgas * 1e-3, 13.8817 * rgas * 1e-3, 13.8817 * rgas * 1e-3};
    const double sig[3] = {3.115, 0.673, 0.673};
    for (int i = 0; i < 6; i++) m->std_q[i] = q[i];
    for (int i = 0; i < 3; i++) {
      m->std_eps[i] = eps[i];
      m->std_sig[i] = sig[i];
    }
    // StdForce coefficients (StdInteraction.F90:206-221): 25 Coulomb terms
    // over H, H, M, L, L; LJ on O-O and on the O-H, H-O, H-H pairs
    const double KX2X = 1.0 / 1e-3, J2I = 1.0 / 10.0;
    m->ncoul = 25;
    for (int t = 0; t < 25; t++) {
      const int a = 1 + t / 5, b = 1 + t % 5;
      m->ci[t] = a;
      m->cj[t] = b;
      m->qq[t] = q[a] * q[b] * coeff() * KX2X * J2I * 1.0;
    }
    auto lj = [&](int a, int b, double *A, double *B) {
      const double ljeps = sqrt(eps[a] * eps[b]);
      const double s6 = ipow((sig[a] + sig[b]) * 0.5, 6);
      *A = 4.0 * ljeps * s6 * s6 * KX2X * J2I * 1.0;
      *B = -4.0 * ljeps * s6 * KX2X * J2I * 1.0;
    };
    lj(0, 0, &m->lj_a, &m->lj_b);
    // the pairs of Traits<G6_NVDE>::xa/xb: (0,1) (0,2) (1,0) (2,0) (1,1) (1,2) (2,1) (2,2)
    const int la[8] = {0, 0, 1, 2, 1, 1, 2, 2}, lb[8] = {1, 2, 0, 0, 1, 2, 1, 2};
    m->nlx = 8;
    for (int t = 0; t < 8; t++) lj(la[t], lb[t], &m->lx_a[t], &m->lx_b[t]);
    set_inverses(m);
    return true;
  }
  if (model == G6_OPLSMEOH) {
    // Rigid_OPLSMeOH_Constructor (RigidModule.F90:391-426): Me, O, H, CoM in
    // the plane z = 0, the coordinates as given (no shift)
    const double ms[4] = {15.0, 16.0, 1.0, 0.0};
    const double x[4] = {0.0133, -0.0634, 0.8152, 0.0};
    const double y[4] = {0.7666, -0.6559, -1.0042, 0.0};
    m->nsite = 4;
    m->Ixx = m->Iyy = m->Izz = m->mass = 0.0;
    for (int i = 0; i < 4; i++) {
      m->bx[i] = x[i];
      m->by[i] = y[i];
      m->bz[i] = 0.0;
      m->Ixx = m->Ixx + ms[i] * (y[i] * y[i] + 0.0 * 0.0);
      m->Iyy = m->Iyy + ms[i] * (0.0 * 0.0 + x[i] * x[i]);
      m->Izz = m->Izz + ms[i] * (x[i] * x[i] + y[i] * y[i]);
      m->mass = m->mass + ms[i];
    }
    // oplsmeoh_setinteraction (OPLSMeOHModule.F90:82-101), the parameters of
    // OPLSMeOHModule.F90:41-49 (eps in kcal/mol times CA = 4.184,
    // PhysConst.F90:29-31)
    const double CA = 4.184;
    const double q[3] = {0.265, -0.700, 0.435};
    const double eps[2] = {0.207 * CA, 0.170 * CA}, sig[2] = {3.775, 3.070};
    for (int i = 0; i < 3; i++) m->std_q[i] = q[i];
    for (int i = 0; i < 2; i++) {
      m->std_eps[i] = eps[i];
      m->std_sig[i] = sig[i];
    }
    // StdForce coefficients (StdInteraction.F90:206-221): 9 Coulomb terms
    // over Me, O, H; LJ on Me-Me and on the Me-O, O-Me, O-O pairs
    const double KX2X = 1.0 / 1e-3, J2I = 1.0 / 10.0;
    m->ncoul = 9;
    for (int t = 0; t < 9; t++) {
      const int a = t / 3, b = t % 3;
      m->ci[t] = a;
      m->cj[t] = b;
      m->qq[t] = q[a] * q[b] * coeff() * KX2X * J2I * 1.0;
    }
    auto lj = [&](int a, int b, double *A, double *B) {
      const double ljeps = sqrt(eps[a] * eps[b]);
      const double s6 = ipow((sig[a] + sig[b]) * 0.5, 6);
      *A = 4.0 * ljeps * s6 * s6 * KX2X * J2I * 1.0;
      *B = -4.0 * ljeps * s6 * KX2X * J2I * 1.0;
    };
    lj(0, 0, &m->lj_a, &m->lj_b);
    // the pairs of Traits<G6_OPLSMEOH>::xa/xb: (0,1) (1,0) (1,1)
    const int la[3] = {0, 1, 1}, lb[3] = {1, 0, 1};
    m->nlx = 3;
    for (int t = 0; t < 3; t++) lj(la[t], lb[t], &m->lx_a[t], &m->lx_b[t]);
    set_inverses(m);
    return true;
  }
  return false;
}

// half-shell stencil, maxradius = 0 branch (NeighborCell_ListHomo,
// GridModule.F90:279-374); offsets packed ox | oy<<8 | oz<<16
static void make_stencil(const int (&nd)[3], Geom *g) {
  int nmax[3], nmin[3];
  for (int a = 0; a < 3; a++) {
    nmax[a] = 1;
    nmin[a] = -1;
    if (nmax[a] > (nd[a] - 1) / 2) nmax[a] = (nd[a] - 1) / 2;
    if (nmin[a] < -(nd[a] / 2)) nmin[a] = -(nd[a] / 2);
  }
  const int nx = nd[0], ny = nd[1], nz = nd[2];
  std::vector<int> mark((size_t)nx * ny * nz, 0);
  auto MK = [&](int x, int y, int z) -> int & { return mark[((size_t)z * ny + y) * nx + x]; };
  g->nst = g->nst0 = 0;
  for (int jz = nmin[2]; jz <= nmax[2]; jz++) {
    const int pz = jz < 0 ? jz + nz : jz, mz = pz == 0 ? 0 : nz - pz;
    for (int jy = nmin[1]; jy <= nmax[1]; jy++) {
      const int py = jy < 0 ? jy + ny : jy, my = py == 0 ? 0 : ny - py;
      for (int jx = nmin[0]; jx <= nmax[0]; jx++) {
        const int px = jx < 0 ? jx + nx : jx, mx = px == 0 ? 0 : nx - px;
        if (MK(px, py, pz) == 0) {
          MK(mx, my, mz) += 1;
          MK(px, py, pz) += 1;
          const int packed = px | (py << 8) | (pz << 16);
          if (MK(px, py, pz) != 1) g->st0[g->nst0++] = packed;
          else g->st[g->nst++] = packed;
        }
      }
    }
  }
}

// ---- MT19937 (Matsumoto & Nishimura 1998; init_genrand of 2002) ---------
struct MT {
  uint32_t s[624];
  int k = 624;
  explicit MT(uint32_t seed) {
    s[0] = seed;
    for (int i = 1; i < 624; i++) s[i] = 1812433253u * (s[i - 1] ^ (s[i - 1] >> 30)) + (uint32_t)i;
  }
  uint32_t next() {
    if (k >= 624) {
      for (int i = 0; i < 624; i++) {
        const uint32_t y = (s[i] & 0x80000000u) | (s[(i + 1) % 624] & 0x7fffffffu);
        s[i] = s[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
      }
      k = 0;
    }
    uint32_t y = s[k++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  // 31-bit uniform in [0,1) as mtrndv (mtrnd.f:280-306)
  double uniform() { return (double)(next() >> 1) * (1.0 / 2147483648.0); }
  // Random_GetNext returns 1 - x (RandomModule.F90:123-130)
  double getnext() { return 1.0 - uniform(); }
};

static std::mutex g_mu;
static std::vector<System *> g_sys;

static System *get(g6_handle h) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (h <= 0 || h > (g6_handle)g_sys.size()) return nullptr;
  return g_sys[h - 1];
}

System *system_of(g6_handle h) { return get(h); }

static int env_int(const char *name, int dflt) {
  const char *e = getenv(name);
  return e ? atoi(e) : dflt;
}

// The MD grid for the rows of both directions: cells of edge >= (rc+skin)/2.
static void set_row_grid(System &s) {
  Geom &m = s.geo_md;
  const double rl = s.rc + s.skin;
  for (int a = 0; a < 3; a++) {
    m.nd[a] = (int)(s.Lb[a] / (0.5 * rl));
    if (m.nd[a] < 1) m.nd[a] = 1;
    m.cellinv[a] = 1.0 / (s.Lb[a] / m.nd[a]);
  }
  m.ncell = m.nd[0] * m.nd[1] * m.nd[2];
  const double percell = (double)s.n / m.ncell;
  s.md.cell_cap = ((int)(2.0 * percell + 33.0) + 31) / 32 * 32;
  if (s.md.cell_cap > s.n + 1) s.md.cell_cap = s.n + 1;
  // G6_CELL_CAP (at g6_init) lowers the capacity, to exercise its error path
  const int cc = env_int("G6_CELL_CAP", 0);
  if (cc > 1 && cc < s.md.cell_cap) s.md.cell_cap = cc;
  s.nblk_md = (s.n + 3) / 4;
  s.md.part = nullptr;
  s.tiles = false;
}

// The MD grid of the tile kernel: cells of edge >= rc + skin, one workgroup
// per (cell, slice).
static void set_tile_grid(System &s) {
  Geom &m = s.geo_md;
  for (int a = 0; a < 3; a++) {
    m.nd[a] = s.tile_ndiv;
    m.cellinv[a] = 1.0 / (s.Lb[a] / m.nd[a]);
  }
  m.ncell = m.nd[0] * m.nd[1] * m.nd[2];
  s.md.cell_cap = s.tile_cap;
  s.nblk_md = m.ncell * s.md.slices;
  s.tiles = true;
}


// Whether a system runs the tile kernel, and its sizes.  The kernel needs
// FP32 pairs, at least 3 cells of edge >= rc + skin per axis (27 distinct
// stencil cells), slots that fit 11 bits and the LDS of the device.
// G6_TILES=0 keeps the rows of both directions; G6_TILE_SLICES and
// G6_TILE_WAVES override the workgroup shape.
static void tile_plan(System &s, double nb_full) {
  s.tile_ok = false;
  if (s.prec != G6_FP32 || env_int("G6_TILES", 1) == 0 || !is_cubic(s)) return;
  const double rl = s.rc + s.skin;
  const int nd = (int)(s.L / rl);
  if (nd < 3) return;
  const int ncell = nd * nd * nd;
  // below one tile per CU the rows of both directions are faster (4096
  // TIP4P at rc 9 A: 64 tiles, 11.7 us against 17.8 us per force evaluation)
  if (ncell < env_int("G6_TILE_MIN", 256)) return;
  const double percell = (double)s.n / ncell;
  int cap = ((int)(2.0 * percell + 33.0) + 31) / 32 * 32;
  if (cap > s.n + 1) cap = s.n + 1;
  if (cap > 1024) return;  // the sort scratch of a wave holds 1024 ids
  // neighbourhood slots: at least the 14 cells' mean plus 15 % and 48 (a
  // liquid's count over 14 cells varies by a few per cent), and as many more
  // as two workgroups per CU leave room for (a lattice start can fill some
  // cells far above the mean); a neighbourhood beyond them is a capacity
  // error, on which g6_step falls back to the rows
  int nbmax = ((int)(1.15 * 14.0 * percell + 48.0) + 15) / 16 * 16;
  // four to eight waves: the step close in block 0 and the block sums need
  // 256 threads
  // Waves per workgroup: a multiple of the CU's four SIMDs, so that every
  // wave gets the same share of its SIMD.  Measured on 65,536 TIP5P at
  // rc 9 A, skin 1 A (1728 tiles): 8 waves, records gathered from HBM, two
  // workgroups per CU 135 us; 4 waves 174 us; 8 waves with the sites staged
  // in LDS (one workgroup per CU) 180 us; the rows of both directions 165 us
  int waves = env_int("G6_TILE_WAVES", 8);
  if (waves < 4 || waves > 8) waves = 8;
  int slices = env_int("G6_TILE_SLICES", 0);
  if (slices < 1) {
    slices = (256 + ncell - 1) / ncell;
    if (slices > 4) slices = 4;
  }
  int lds_max = 0;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, 0) != hipSuccess)
    return;
  const int nq = s.mt.nsite - 1;
  // LDS besides the dynamic part: the kernel's static arrays and the close
  const int fixed = 2048;
  const bool stage = env_int("G6_TILE_STAGE", 0) != 0;
  {
    // the most slots that two workgroups per CU leave room for
    int room = std::min(14 * (cap - 1), (1 << 11) - 1) / 16 * 16;
    while (room > nbmax && tile_layout(nq, room, cap, waves, stage).bytes + fixed > lds_max / 2)
      room -= 16;
    if (room > nbmax) nbmax = room;
  }
  if (nbmax > 14 * (cap - 1)) nbmax = 14 * (cap - 1);
  if (nbmax > (1 << 11) - 1) nbmax = (1 << 11) - 1 - 15;
  // G6_TILE_SLOTS (at g6_init) lowers the slots, to exercise the fallback
  const int ts = env_int("G6_TILE_SLOTS", 0);
  if (ts >= 16 && ts < nbmax) nbmax = ts / 16 * 16;
  TileLayout lay = tile_layout(nq, nbmax, cap, waves, stage);
  while (lay.bytes + fixed > lds_max && waves > 4) {
    waves -= 1;
    lay = tile_layout(nq, nbmax, cap, waves, stage);
  }
  if (lay.bytes + fixed > lds_max) return;
  lay.xcd = env_int("G6_TILE_XCD", 1) != 0 ? 1 : 0;
  s.tile_ndiv = nd;
  s.tile_cap = cap;
  s.tile_waves = waves;
  s.tile_lay = lay;
  s.md.nbmax = nbmax;
  s.md.slices = slices;
  s.md.nds = 14 * slices;
  int mr = ((int)(nb_full + 64.0) + 63) / 64 * 64;
  if (mr > s.n) mr = s.n;
  s.md.maxrow = mr;
  s.tile_ok = true;
}

static int alloc_all(System &s) {
  const int n = s.n, ns = s.ns, nc = s.geo.ncell, ncm = s.geo_md.ncell;
  auto A = [](void **p, size_t b) { return hipMalloc(p, b > 0 ? b : 8); };
#define AL(ptr, bytes) G6_CHECK(A((void **)&(ptr), (bytes)))
  MdBuf &m = s.md;
  AL(m.com, sizeof(double) * 3 * n);
  AL(m.dx, sizeof(double) * 15 * n);
  AL(m.w, sizeof(double) * 15 * n);
  AL(m.quat, sizeof(double) * 20 * n);
  AL(m.tmat, sizeof(double) * 9 * n);
  AL(m.rec, sizeof(double) * (size_t)m.rs * n);
  AL(m.com_ref, sizeof(double) * 3 * n);
  AL(m.cell, sizeof(int) * n);
  // cell lists sized for the row grid and, when the system qualifies, the tile grid
  size_t ncap = (size_t)ncm * m.cell_cap;
  if (s.tile_ok) {
    const size_t nt = (size_t)s.tile_ndiv * s.tile_ndiv * s.tile_ndiv * s.tile_cap;
    if (nt > ncap) ncap = nt;
  }
  AL(m.cellres, sizeof(int) * 2 * ncap);
  AL(m.nbr_j, sizeof(int) * (size_t)n * m.maxnb);
  AL(m.nbr_code, sizeof(int) * (size_t)n * m.maxnb);
  AL(m.nbr_cnt, sizeof(int) * n);
  {
    size_t nb = (size_t)s.nblk_md;
    if (s.tile_ok) {
      const size_t nt = (size_t)s.tile_ndiv * s.tile_ndiv * s.tile_ndiv * s.md.slices;
      if (nt > nb) nb = nt;
    }
    AL(m.blockpart, sizeof(double) * 2 * kBP * nb);
  }
  AL(m.blockek, sizeof(double) * kEkLen * (size_t)pre_waves(n));
  AL(m.frc, sizeof(double) * 6 * n);
  AL(m.nose, sizeof(double) * 9);
  AL(m.thermo, sizeof(double) * G6_THERMO_LEN);
  AL(m.flags, sizeof(int) * 8);
  s.nose = m.nose;
  s.thermo = m.thermo;
  s.flags = m.flags;
  // reference-list entry point (g6_pairlist)
  AL(s.cell, sizeof(int) * n);
  AL(s.cell_count, sizeof(int) * (nc + 1));
  AL(s.cell_start, sizeof(int) * (nc + 1));
  AL(s.cell_fill, sizeof(int) * (nc + 1));
  AL(s.sorted_tmp, sizeof(int) * n);
  AL(s.sorted, sizeof(int) * n);
  AL(s.cnt, sizeof(int) * 2 * n);
  AL(s.off, sizeof(int) * (2 * n + 1));
  AL(s.pi, sizeof(int) * (size_t)s.pcap);
  AL(s.pj, sizeof(int) * (size_t)s.pcap);
  AL(s.scode, sizeof(int) * (size_t)s.pcap);
  AL(s.com_ref, sizeof(double) * 3 * n);
  AL(s.lflags, sizeof(int) * 8);
#undef AL
  G6_CHECK(hipMemset(m.com, 0, sizeof(double) * 3 * n));
  G6_CHECK(hipMemset(m.dx, 0, sizeof(double) * 15 * n));
  G6_CHECK(hipMemset(m.w, 0, sizeof(double) * 15 * n));
  G6_CHECK(hipMemset(m.quat, 0, sizeof(double) * 20 * n));
  G6_CHECK(hipMemset(m.nbr_cnt, 0, sizeof(int) * n));
  G6_CHECK(hipMemset(m.com_ref, 0, sizeof(double) * 3 * n));
  G6_CHECK(hipMemset(m.nose, 0, sizeof(double) * 9));
  G6_CHECK(hipMemset(m.thermo, 0, sizeof(double) * G6_THERMO_LEN));
  return 0;
}

// the tile kernel's buffers (the system qualified in tile_plan)
static int tile_alloc(System &s) {
  MdBuf &m = s.md;
  const size_t nt = (size_t)s.tile_ndiv * s.tile_ndiv * s.tile_ndiv;
  auto A = [](void **p, size_t b) { return hipMalloc(p, b > 0 ? b : 8); };
  G6_CHECK(A((void **)&m.part, sizeof(float) * 6 * (size_t)m.nds * s.n));
  G6_CHECK(A((void **)&m.nb_id, sizeof(int) * nt * m.nbmax));
  G6_CHECK(A((void **)&m.nb_base, sizeof(int) * 16 * nt));
  G6_CHECK(A((void **)&m.trow, sizeof(unsigned) * (size_t)m.maxrow * s.n));
  G6_CHECK(A((void **)&m.trow_cnt, sizeof(int) * s.n));
  G6_CHECK(A((void **)&m.tcnt, sizeof(int) * nt * s.tile_cap));
  G6_CHECK(hipMemset(m.part, 0, sizeof(float) * 6 * (size_t)m.nds * s.n));
  if (s.tile_lay.bytes > 64 * 1024) G6_CHECK(tile_set_lds_limit(s.tile_lay.bytes) == 0 ? hipSuccess : hipErrorInvalidValue);
  s.tile_part = m.part;
  return 0;
}

// neighbour rows and cell lists ready for a fresh first step
static int reset_step_state(System &s) {
  int fl[8] = {0, 0, 1, 1, 0, 0, 0, 0};
  G6_CHECK(hipMemcpy(s.md.flags, fl, sizeof(fl), hipMemcpyHostToDevice));
  G6_CHECK(hipMemset(s.md.cellres, 0, sizeof(int) * 2 * (size_t)s.geo_md.ncell * s.md.cell_cap));
  return 0;
}

int seam_reserve(System &s, size_t bytes, char **base) {
  if (bytes > s.seam_cap) {
    G6_CHECK(hipStreamSynchronize(s.stream));
    if (s.seam) (void)hipFree(s.seam);
    s.seam = nullptr;
    s.seam_cap = 0;
    G6_CHECK(hipMalloc((void **)&s.seam, bytes));
    s.seam_cap = bytes;
  }
  *base = s.seam;
  return 0;
}

// Back to the tiles after a capacity fallback (a lattice start can fill
// some tile neighbourhoods far above a liquid's counts for a few steps).
static int use_tile_grid(System &s) {
  if (s.tiles || !s.tile_ok) return 0;
  G6_CHECK(hipStreamSynchronize(s.stream));
  s.md.part = s.tile_part;
  set_tile_grid(s);
  drop_graphs(s);
  return reset_step_state(s);
}

// a capacity error moved a tile system to the rows: when to try the tiles
// again (G6_TILE_RETRY steps later, then twice as long after each failure)
static void note_tile_exit(System &s) {
  s.tile_exit = s.last_flag;
  if (s.tile_retry_gap == 0) s.tile_retry_gap = std::max(1, env_int("G6_TILE_RETRY", 1000));
  s.tile_retry_at = s.steps_done + s.tile_retry_gap;
  s.tile_retry_gap = (int)std::min(2LL * s.tile_retry_gap, 1LL << 24);
}

// Back to the rows of both directions (ladders and the barostat use them).
int use_row_grid(System &s) {
  if (!s.tiles) return 0;
  G6_CHECK(hipStreamSynchronize(s.stream));
  set_row_grid(s);
  drop_graphs(s);
  return reset_step_state(s);
}

static void free_all(System &s) {
  const MdBuf &m = s.md;
  void *ps[] = {m.com, m.dx, m.w, m.quat, m.tmat, m.rec, m.com_ref,
                m.cell, m.cellres, m.nbr_j, m.nbr_code, m.nbr_cnt, m.blockpart,
                m.blockek, m.frc, m.nose, m.thermo, m.flags, s.cell, s.cell_count, s.cell_start,
                s.cell_fill, s.sorted_tmp, s.sorted, s.cnt, s.off, s.pi, s.pj, s.scode,
                s.com_ref, s.lflags, m.box, m.anders, s.tile_part, m.nb_id, m.nb_base,
                m.trow, m.trow_cnt, m.tcnt, s.snap, s.seam};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  mc_free(s);
  drop_graphs(s);
  for (auto &e : s.ev) (void)hipEventDestroy(e);
  if (s.stream) (void)hipStreamDestroy(s.stream);
}

// k steps: predict, forces; (correct + predict, forces) x (k-1); correct.
// Each call is self-contained, so a captured graph of k steps can be replayed.
// With the Andersen barostat the box of step t follows from the pressure of
// step t-1, which needs every molecule's corrector: correct, close (thermo,
// Nose and Andersen correct and predict, the new box), predict and bin in
// that box, forces; four kernels per step.
static void launch_steps(System &s, int k) {
  if (s.andersen) {
    for (int q = 0; q < k; q++) {
      const int p = q & 1;
      if (q > 0) md_launch_pre(s, true, false, false, 0);
      // clears the cells and rebuild flag of parity p before they are filled
      md_launch_close(s, (q > 0 ? 1 : 0) | 2 | 4, 1 - p, 1 - p);
      md_launch_pre(s, false, true, true, p);
      md_launch_main(s, nullptr, p, 0);
    }
    md_launch_pre(s, true, false, false, 0);
    md_launch_close(s, 1 | 8, 0, (k - 1) & 1);
    return;
  }
  for (int q = 0; q < k; q++) {
    const int p = q & 1;
    md_launch_pre(s, q > 0, true, true, p);
    if (q == 0) md_launch_close(s, 2 | 4, p, 1 - p);
    md_launch_main(s, nullptr, p, q > 0 ? (1 | 2 | 4) : 0);
  }
  md_launch_pre(s, true, false, false, 0);
  md_launch_close(s, 1 | 8, 0, (k - 1) & 1);
}

void drop_graphs(System &s) {
  s.graph_gen++;
  if (s.graph) (void)hipGraphExecDestroy(s.graph);
  s.graph = nullptr;
  for (int q = 0; q < 4; q++) {
    if (s.graph_k[q]) (void)hipGraphExecDestroy(s.graph_k[q]);
    s.graph_k[q] = nullptr;
    s.graph_k_steps[q] = 0;
  }
}

// steps per captured hipGraph (G6_GRAPH_STEPS; 0 disables graphs)
int graph_steps() {
  const char *e = getenv("G6_GRAPH_STEPS");
  int v = e ? atoi(e) : 50;
  return v < 0 ? 0 : (v > 1000 ? 1000 : v);
}

}  // namespace g6

using namespace g6;

extern "C" {

int g6_model_nsite(int model) {
  ModelTable m;
  return make_model_table(model, 0, &m) ? m.nsite : 0;
}

static int scratch_address(void **p);

static int init_system(g6_handle *h, int model, int n, const double *box3, double rc,
                       double taper, const double *voxa3, double dt, int prec, double skin) {
  if (!h || n < 2 || !(rc > 0) || !(taper >= 0) || taper > rc || !(dt > 0) ||
      (prec != G6_FP64 && prec != G6_FP32) || !(skin >= 0))
    return G6_EARG;
  for (int a = 0; a < 3; a++)
    if (!(box3[a] > 0) || !(voxa3[a] > 0)) return G6_EARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return G6_ENODEVICE;
  System *s = new System();
  if (!make_model_table(model, prec, &s->mt)) {
    delete s;
    return G6_EARG;
  }
  s->model = model;
  s->fast = 1;
  s->kmodel = kernel_model(model, true);
  s->n = n;
  s->ns = s->mt.nsite;
  s->prec = prec;
  for (int a = 0; a < 3; a++) s->Lb[a] = box3[a];
  s->L = box3[0];
  s->rc = rc;
  s->taper = taper;
  s->req = voxa3[0];
  s->dt = dt;
  s->skin = skin;
  Geom &g = s->geo;
  g.n = n;
  for (int a = 0; a < 3; a++) {
    // Grid_ResetDivision: integer part of L/req per axis (GridModule.F90:494-499)
    g.nd[a] = (int)(box3[a] / voxa3[a]);
    if (g.nd[a] < 1) g.nd[a] = 1;
    if (rc + skin > 0.5 * box3[a] * (1.0 + 1e-12) && g.nd[a] > 2) {
      delete s;  // the half-shell stencil of 27 cells needs rc + skin <= L/2
      return G6_EARG;
    }
    g.L[a] = box3[a];
    g.invL[a] = 1.0 / box3[a];                 // Box_ReadBXLA / Box_ReadBOX3
    g.cellinv[a] = 1.0 / (box3[a] / g.nd[a]);  // Grid_ResetCellSize + invert
  }
  g.ncell = g.nd[0] * g.nd[1] * g.nd[2];
  make_stencil(g.nd, &g);
  const double vol = (box3[0] * box3[1]) * box3[2];
  const double rl = rc + skin;
  g.rr = rl * rl;
  g.rc2 = rc * rc;
  g.rin = rc - taper;
  g.rout = rc;
  g.c1 = taper > 0 ? 1.0 / ipow(g.rin - g.rout, 5) : 0.0;  // cutoff_initialize
  g.c2 = g.c1 * 30.0;
  // pair capacity: 1.6x the ideal-gas half-shell count within rc+skin, +64/molecule
  const double per = 0.5 * 4.0 / 3.0 * kPI * rl * rl * rl * n / vol;
  const double capd = (double)n * (per * 1.6 + 64.0);
  const double all = 0.5 * (double)n * (double)(n - 1);
  s->pcap = (int)fmin(fmin(capd, all + 64.0), 2.0e9);
  // the MD grid: cells no smaller than (rc + skin)/2, searched with a 5^3
  // stencil pruned per molecule (all cells once when ndiv < 5); tile systems
  // switch to cells no smaller than rc + skin below
  {
    Geom &m = s->geo_md;
    m = g;
    MdBuf &b = s->md;
    b.rs = prec == G6_FP32 ? kRecF : kRecD;
    set_row_grid(*s);
    const double rho = n / vol;
    const double nb = 4.0 / 3.0 * kPI * rl * rl * rl * rho;
    b.maxnb = ((int)(1.5 * nb + 64.0) + 63) / 64 * 64;
    if (b.maxnb > n) b.maxnb = n;
    b.half_skin2 = skin > 0 ? 0.25 * skin * skin : 0.0;
    s->pkf = make_pair_k<float>(s->mt, m);
    s->pkd = make_pair_k<double>(s->mt, m);
    // the force kernel keeps four rows in LDS (64 KiB of dynamic LDS at most)
    const int rb = main_region_bytes(b.maxnb, prec == G6_FP32 ? 4 : 8);
    if (4 * rb > 60 * 1024) {
      delete s;
      return G6_EARG;
    }
    tile_plan(*s, nb);
  }
  int rcode = alloc_all(*s);
  if (rcode == 0 && s->tile_ok) {
    rcode = tile_alloc(*s);
    if (rcode == 0) set_tile_grid(*s);
  }
  if (rcode == 0 && hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    rcode = G6_EHIP;
  for (int k = 0; rcode == 0 && k < 8; k++)
    if (hipEventCreate(&s->ev[k]) != hipSuccess) rcode = G6_EHIP;
  if (rcode == 0) rcode = reset_step_state(*s);
  if (rcode == 0) {
    // the no-allocation entry point's device array, on this device
    void *scr = nullptr;
    rcode = scratch_address(&scr);
  }
  if (rcode != 0) {
    free_all(*s);
    delete s;
    return rcode;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_sys.push_back(s);
  *h = (g6_handle)g_sys.size();
  return 0;
}

int g6_init(g6_handle *h, int model, int n, double box_l, double rc, double taper,
            double voxa, double dt, int prec, double skin) {
  const double box3[3] = {box_l, box_l, box_l}, voxa3[3] = {voxa, voxa, voxa};
  return init_system(h, model, n, box3, rc, taper, voxa3, dt, prec, skin);
}

int g6_init_box3(g6_handle *h, int model, int n, const double *box3, double rc, double taper,
                 const double *voxa3, double dt, int prec, double skin) {
  if (!box3 || !voxa3) return G6_EARG;
  return init_system(h, model, n, box3, rc, taper, voxa3, dt, prec, skin);
}

int g6_finalize(g6_handle h) {
  System *s = get(h);
  if (!s) return G6_EARG;
  (void)hipStreamSynchronize(s->stream);
  npt_forget(s);
  free_all(*s);
  delete s;
  std::lock_guard<std::mutex> lk(g_mu);
  g_sys[h - 1] = nullptr;
  return 0;
}

int g6_upload_state(g6_handle h, const double *com, const double *quat, const double *dx,
                    const double *w) {
  System *s = get(h);
  if (!s || !com || !quat || !dx || !w) return G6_EARG;
  const int n = s->n;
  G6_CHECK(hipMemcpy(s->md.com, com, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  G6_CHECK(hipMemcpy(s->md.quat, quat, sizeof(double) * 20 * n, hipMemcpyHostToDevice));
  G6_CHECK(hipMemcpy(s->md.dx, dx, sizeof(double) * 15 * n, hipMemcpyHostToDevice));
  G6_CHECK(hipMemcpy(s->md.w, w, sizeof(double) * 15 * n, hipMemcpyHostToDevice));
  return reset_step_state(*s);
}

int g6_download_state(g6_handle h, double *com, double *quat, double *dx, double *w) {
  System *s = get(h);
  if (!s) return G6_EARG;
  const int n = s->n;
  G6_CHECK(hipStreamSynchronize(s->stream));
  if (com) G6_CHECK(hipMemcpy(com, s->md.com, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  if (quat) G6_CHECK(hipMemcpy(quat, s->md.quat, sizeof(double) * 20 * n, hipMemcpyDeviceToHost));
  if (dx) G6_CHECK(hipMemcpy(dx, s->md.dx, sizeof(double) * 15 * n, hipMemcpyDeviceToHost));
  if (w) G6_CHECK(hipMemcpy(w, s->md.w, sizeof(double) * 15 * n, hipMemcpyDeviceToHost));
  return 0;
}

int g6_set_nose(g6_handle h, int active, const double *nose) {
  System *s = get(h);
  if (!s) return G6_EARG;
  if (active && (!nose || !(nose[1] > 0))) return G6_EARG;
  s->nose_active = active != 0;
  double z[9] = {0};
  if (nose) memcpy(z, nose, sizeof(z));
  G6_CHECK(hipStreamSynchronize(s->stream));
  G6_CHECK(hipMemcpy(s->nose, z, sizeof(z), hipMemcpyHostToDevice));
  drop_graphs(*s);  // the captured launches embed the Nose flag
  return 0;
}

int g6_set_fast(g6_handle h, int fast) {
  System *s = get(h);
  if (!s) return G6_EARG;
  ModelTable m;
  if (!make_model_table(s->model, s->prec, &m, fast != 0)) return G6_EARG;
  G6_CHECK(hipStreamSynchronize(s->stream));
  s->mt = m;
  s->fast = fast != 0;
  s->kmodel = kernel_model(s->model, s->fast);
  s->pkf = make_pair_k<float>(s->mt, s->geo_md);
  s->pkd = make_pair_k<double>(s->mt, s->geo_md);
  drop_graphs(*s);  // the captured launches embed the pair constants
  return 0;
}

int g6_get_nose(g6_handle h, double *nose) {
  System *s = get(h);
  if (!s || !nose) return G6_EARG;
  G6_CHECK(hipStreamSynchronize(s->stream));
  G6_CHECK(hipMemcpy(nose, s->nose, sizeof(double) * 9, hipMemcpyDeviceToHost));
  return 0;
}

int g6_vreset(g6_handle h, double temp, uint32_t seed) {
  System *s = get(h);
  if (!s || !(temp >= 0)) return G6_EARG;
  return vreset_systems(&s, 1, temp, seed);
}

}  // extern "C"

namespace g6 {

int vreset_systems(System *const *sys, int nsys, double temp, uint32_t seed) {
  // Rigid_VReset2 (RigidModule.F90:1016-1053) with Random_Normal
  // (RandomModule.F90:28-36): dx(1) = N(0, T*rgas*J2I/m)*dt, w(1) = N(0, T*rgas*J2I/I);
  // the components of a mixture draw in turn from one stream (Mixture.F90:1336-1346)
  MT mt(seed);
  for (int c = 0; c < nsys; c++) {
    System *s = sys[c];
    const int n = s->n;
    std::vector<double> dx(15 * (size_t)n), w(15 * (size_t)n);
    G6_CHECK(hipStreamSynchronize(s->stream));
    G6_CHECK(hipMemcpy(dx.data(), s->md.dx, sizeof(double) * 15 * n, hipMemcpyDeviceToHost));
    G6_CHECK(hipMemcpy(w.data(), s->md.w, sizeof(double) * 15 * n, hipMemcpyDeviceToHost));
    const double rgas = kNA * kKB, J2I = 1.0 / 10.0;
    const double massi = 1.0 / s->mt.mass;
    const double ii[3] = {1.0 / s->mt.Ixx, 1.0 / s->mt.Iyy, 1.0 / s->mt.Izz};
    auto normal = [](double var, double r1, double r2) {
      const double r = sqrt(-2.0 * log(r1));
      const double theta = 2.0 * kPI * r2;
      return r * sin(theta) * sqrt(var) + 0.0;
    };
    for (int i = 0; i < n; i++) {
      for (int a = 0; a < 3; a++) {
        const double r1 = mt.getnext(), r2 = mt.getnext();
        dx[(size_t)a * n + i] = normal(temp * rgas * J2I * massi, r1, r2) * s->dt;
      }
      for (int a = 0; a < 3; a++) {
        const double r1 = mt.getnext(), r2 = mt.getnext();
        w[(size_t)a * n + i] = normal(temp * rgas * J2I * ii[a], r1, r2);
      }
    }
    G6_CHECK(hipMemcpy(s->md.dx, dx.data(), sizeof(double) * 15 * n, hipMemcpyHostToDevice));
    G6_CHECK(hipMemcpy(s->md.w, w.data(), sizeof(double) * 15 * n, hipMemcpyHostToDevice));
  }
  return 0;
}

}  // namespace g6

extern "C" {

// flags[0]: 1 pair capacity, 2 image shift beyond 511 box lengths,
// 3 cell capacity, 4 neighbour-row capacity
static int check_flags(System &s) {
  int fl[2] = {0, 0};
  G6_CHECK(hipMemcpyAsync(fl, s.md.flags, sizeof(fl), hipMemcpyDeviceToHost, s.stream));
  G6_CHECK(hipStreamSynchronize(s.stream));
  s.last_flag = fl[0];
  if (fl[0] == 1 || fl[0] == 3 || fl[0] == 4 || fl[0] == 5 || fl[0] == 6) return G6_ECAPACITY;
  if (fl[0] == 2) return G6_EARG;
  return 0;
}

// Device copy of the state a g6_step call starts from (Nordsieck arrays,
// the Nose and barostat variables), so that a call that runs out of a
// capacity leaves the state it was given.
static int snap_state(System &s, bool restore) {
  const size_t n = (size_t)s.n;
  if (!s.snap) G6_CHECK(hipMalloc((void **)&s.snap, sizeof(double) * (53 * n + 25)));
  struct Part {
    double *d;
    size_t len;
  } parts[] = {{s.md.com, 3 * n}, {s.md.dx, 15 * n}, {s.md.w, 15 * n},   {s.md.quat, 20 * n},
               {s.md.nose, 9},    {s.md.box, 8},      {s.md.anders, 8}};
  double *p = s.snap;
  for (const Part &q : parts) {
    if (!q.d) continue;
    G6_CHECK(hipMemcpyAsync(restore ? q.d : p, restore ? p : q.d, sizeof(double) * q.len,
                            hipMemcpyDeviceToDevice, s.stream));
    p += q.len;
  }
  return 0;
}

static int step_once(System *s, int nsteps);

int g6_step(g6_handle h, int nsteps) {
  System *s = get(h);
  if (!s || nsteps < 0) return G6_EARG;
  if (!s->tiles && s->tile_ok && s->tile_exit != 0 && !s->andersen &&
      s->steps_done >= s->tile_retry_at) {
    const int rt = use_tile_grid(*s);
    if (rt != 0) return rt;
  }
  const int sv = snap_state(*s, false);
  if (sv != 0) return sv;
  const int rc = step_once(s, nsteps);
  if (rc == 0 || rc == G6_EHIP) return rc;
  // a capacity ran out (or an image shift left its code): back to the state
  // of the call's start; a tile system reruns the call on the rows of both
  // directions, whose capacities are its own, and keeps them
  const int rs = snap_state(*s, true);
  if (rs != 0) return rs;
  G6_CHECK(hipStreamSynchronize(s->stream));
  s->steps_done -= nsteps;
  if (s->tiles) {
    note_tile_exit(*s);
    const int rg = use_row_grid(*s);
    if (rg != 0) return rg;
    return g6_step(h, nsteps);
  }
  const int rr = reset_step_state(*s);
  return rr != 0 ? rr : rc;
}

static int capture_steps(System &s, int k, hipGraphExec_t *g) {
  hipGraph_t gr;
  G6_CHECK(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
  launch_steps(s, k);
  G6_CHECK(hipStreamEndCapture(s.stream, &gr));
  const hipError_t e = hipGraphInstantiate(g, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  G6_CHECK(e);
  return 0;
}

// G-step graphs replayed, then a graph of the remaining count (cached by
// count, four at most), so that every call replays graphs
static int step_once(System *s, int nsteps) {
  const int G = graph_steps();
  if (G == 0) {
    launch_steps(*s, nsteps);
  } else {
    if (nsteps >= G) {
      if (!s->graph || s->graph_steps != G) {
        if (s->graph) (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
        const int rc = capture_steps(*s, G, &s->graph);
        if (rc != 0) return rc;
        s->graph_steps = G;
      }
      for (int k = 0; k + G <= nsteps; k += G) G6_CHECK(hipGraphLaunch(s->graph, s->stream));
    }
    const int rest = nsteps % G;
    if (rest > 0) {
      int slot = -1;
      for (int q = 0; q < 4; q++)
        if (s->graph_k[q] && s->graph_k_steps[q] == rest) slot = q;
      if (slot < 0) {
        slot = s->graph_k_next;
        s->graph_k_next = (slot + 1) % 4;
        if (s->graph_k[slot]) (void)hipGraphExecDestroy(s->graph_k[slot]);
        s->graph_k[slot] = nullptr;
        const int rc = capture_steps(*s, rest, &s->graph_k[slot]);
        if (rc != 0) return rc;
        s->graph_k_steps[slot] = rest;
      }
      G6_CHECK(hipGraphLaunch(s->graph_k[slot], s->stream));
    }
  }
  G6_CHECK(hipGetLastError());
  s->steps_done += nsteps;
  return check_flags(*s);
}

int g6_step_timed(g6_handle h, int nsteps, double *t) {
  System *s = get(h);
  if (!s || nsteps < 1 || !t || s->andersen) return G6_EARG;
  double acc[5] = {0, 0, 0, 0, 0};
  for (int k = 0; k < nsteps; k++) {
    // one step: predict, forces, correct (the corrector and the next
    // predictor share a kernel in g6_step; here they are timed apart)
    G6_CHECK(hipEventRecord(s->ev[0], s->stream));
    md_launch_pre(*s, false, true, true, 0);
    md_launch_close(*s, 2 | 4, 0, 1);
    G6_CHECK(hipEventRecord(s->ev[1], s->stream));
    md_launch_main(*s, nullptr, 0, 0);
    G6_CHECK(hipEventRecord(s->ev[2], s->stream));
    md_launch_pre(*s, true, false, false, 0);
    md_launch_close(*s, 1 | 8, 0, 0);
    G6_CHECK(hipEventRecord(s->ev[3], s->stream));
    G6_CHECK(hipEventSynchronize(s->ev[3]));
    float a, b, c, e;
    G6_CHECK(hipEventElapsedTime(&a, s->ev[0], s->ev[1]));
    G6_CHECK(hipEventElapsedTime(&b, s->ev[1], s->ev[2]));
    G6_CHECK(hipEventElapsedTime(&c, s->ev[2], s->ev[3]));
    G6_CHECK(hipEventElapsedTime(&e, s->ev[0], s->ev[3]));
    acc[0] += b;   // force kernel (rows rebuilt in it when due)
    acc[1] += 0;   // no separate list build
    acc[2] += a + c;
    acc[3] += e;
    double th[G6_THERMO_LEN];
    G6_CHECK(hipMemcpy(th, s->thermo, sizeof(th), hipMemcpyDeviceToHost));
    acc[4] += th[14];
  }
  for (int q = 0; q < 5; q++) t[q] = acc[q] / nsteps;
  s->steps_done += nsteps;
  return check_flags(*s);
}

int g6_time_force(g6_handle h, int nlaunch, double *ms) {
  System *s = get(h);
  if (!s || nlaunch < 1 || nlaunch > 10000 || !ms || s->andersen) return G6_EARG;
  // the force kernel alone, nlaunch times back to back in one graph, on the
  // state a g6_step call leaves: its records and rows are consistent, the
  // rebuild flags clear; each launch writes the same forces and partials
  // (eager launches when G6_GRAPH_STEPS=0, as for kernel tracing)
  hipGraph_t gr;
  hipGraphExec_t ge = nullptr;
  const bool graph = graph_steps() > 0;
  G6_CHECK(hipStreamSynchronize(s->stream));
  if (graph) {
    G6_CHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < nlaunch; k++) md_launch_main(*s, nullptr, 0, 0);
    G6_CHECK(hipStreamEndCapture(s->stream, &gr));
    G6_CHECK(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
    G6_CHECK(hipGraphDestroy(gr));
    G6_CHECK(hipGraphLaunch(ge, s->stream));  // warm
  } else {
    md_launch_main(*s, nullptr, 0, 0);
  }
  G6_CHECK(hipEventRecord(s->ev[4], s->stream));
  if (graph) {
    G6_CHECK(hipGraphLaunch(ge, s->stream));
  } else {
    for (int k = 0; k < nlaunch; k++) md_launch_main(*s, nullptr, 0, 0);
  }
  G6_CHECK(hipEventRecord(s->ev[5], s->stream));
  G6_CHECK(hipEventSynchronize(s->ev[5]));
  float e = 0;
  G6_CHECK(hipEventElapsedTime(&e, s->ev[4], s->ev[5]));
  if (ge) G6_CHECK(hipGraphExecDestroy(ge));
  *ms = e / nlaunch;
  return check_flags(*s);
}

int g6_md_kernel(g6_handle h, int32_t *info) {
  System *s = get(h);
  if (!s || !info) return G6_EARG;
  info[0] = s->tiles ? 1 : 0;
  info[1] = s->tiles ? s->tile_ndiv : 0;
  info[2] = s->tiles ? s->md.slices : 0;
  info[3] = s->tiles ? s->tile_waves : 0;
  info[4] = s->tiles ? s->md.nbmax : 0;
  info[5] = s->tile_exit;
  return 0;
}

// Forces through the device buffer out[6][n] into the host arrays.  The
// rows are rebuilt from the current positions; the state is not moved and the
// thermostat is not advanced.  With fallback, a tile system that runs out of
// a capacity continues on the rows of both directions.
static int eval_forces_into(System *s, double *out, double *force, double *torque,
                            double *thermo, bool fallback) {
  const int n = s->n;
  for (int attempt = 0;; attempt++) {
    md_launch_flag(*s, 2, 1);  // rebuild the rows of parity 0
    md_launch_pre(*s, false, false, true, 0);
    md_launch_main(*s, out, 0, 0);
    md_launch_close(*s, 16 | 8, 0, 0);
    G6_CHECK(hipGetLastError());
    G6_CHECK(hipMemcpyAsync(force, out, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s->stream));
    G6_CHECK(hipMemcpyAsync(torque, out + 3 * n, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                            s->stream));
    G6_CHECK(hipStreamSynchronize(s->stream));
    const int rc = check_flags(*s);
    if (rc != 0 && rc != G6_EHIP && s->tiles && fallback && attempt == 0) {
      note_tile_exit(*s);
      const int rg = use_row_grid(*s);
      if (rg != 0) return rg;
      continue;
    }
    if (rc != 0) return rc;
    if (thermo)
      G6_CHECK(hipMemcpy(thermo, s->thermo, sizeof(double) * G6_THERMO_LEN, hipMemcpyDeviceToHost));
    return 0;
  }
}

int g6_eval_forces(g6_handle h, double *force, double *torque, double *thermo) {
  System *s = get(h);
  if (!s || !force || !torque) return G6_EARG;
  double *out = nullptr;
  G6_CHECK(hipMalloc(&out, sizeof(double) * 6 * s->n));
  const int rc = eval_forces_into(s, out, force, torque, thermo, true);
  (void)hipFree(out);
  return rc;
}

// ---- the entry point that allocates nothing -------------------------------
// g6_eval_forces_noalloc stages force and torque (6 n doubles) in this
// file-scope device array instead of a hipMalloc'd buffer.  Every GPU the
// library runs on holds its own copy until the process ends; host code uses
// the address hipGetSymbolAddress gives on each device, never the array's
// name.  Calls that overlap on one GPU share it and race.
constexpr size_t kScratchBytes = (size_t)16 << 20;  // 6 n doubles for n <= 349,525
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scr_mu;
static void *g_scr_ptr[kMaxDevices];  // scratch's device address, per device

// the address of scratch on the current device, resolved once per device
static int scratch_address(void **p) {
  int dev = 0;
  G6_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return G6_EARG;
  std::lock_guard<std::mutex> lk(g_scr_mu);
  if (!g_scr_ptr[dev]) {
    void *q = nullptr;
    if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return G6_EHIP;
    g_scr_ptr[dev] = q;
  }
  *p = g_scr_ptr[dev];
  return 0;
}

int g6_eval_forces_noalloc(g6_handle h, double *force, double *torque, double *thermo) {
  System *s = get(h);
  if (!s || !force || !torque) return G6_EARG;
  if (sizeof(double) * 6 * (size_t)s->n > kScratchBytes) return G6_ECAPACITY;
  void *p = nullptr;
  const int rc = scratch_address(&p);
  if (rc != 0) return rc;
  // every element the kernels read from the buffer is written first in the
  // same call, so the last call's contents need no clearing
  return eval_forces_into(s, static_cast<double *>(p), force, torque, thermo, false);
}

int g6_set_andersen(g6_handle h, int mode, double pex, double mass, const double *v) {
  System *s = get(h);
  if (!s) return G6_EARG;
  if (mode == 0) {
    if (s->andersen) {
      G6_CHECK(hipStreamSynchronize(s->stream));
      (void)hipFree(s->md.box);
      (void)hipFree(s->md.anders);
      s->md.box = s->md.anders = nullptr;
      s->andersen = false;
    }
  } else {
    // orthorhombic (every edge follows the volume) or orz (the z edge alone);
    // every argument is checked before the handle changes: a failed call
    // leaves the fixed-box path as it was
    if ((mode != 1 && mode != 2) || !(mass > 0)) return G6_EARG;
    if (v && !(v[0] > 0)) return G6_EARG;
    const double dof = 6.0 * s->n;
    const double rl = s->rc + s->skin;
    // the barostat runs on the rows of both directions (its grid has cells
    // of edge >= (rc + skin)/2 in the largest box it allows)
    const int rg = use_row_grid(*s);
    if (rg != 0) return rg;
    s->tile_ok = false;  // the tile grid was planned for the fixed box
    // the box is the handle's (a restart's @BOX3); v continues a saved
    // barostat state (@PCN5), NULL starts one
    double bx[kBoxLen];
    for (int a = 0; a < 3; a++) {
      const double L = s->Lb[a], nd = s->geo_md.nd[a];
      bx[a] = L;
      bx[3 + a] = 1.0 / L;
      bx[6 + a] = 1.0 / (L / nd);
      bx[9 + a] = L;
      bx[12 + a] = nd;
      if (L < nd * (0.5 * rl)) return G6_EARG;
    }
    bx[15] = 0.5 * rl;
    bx[16] = mode;
    // internal mass: per-molecule mass * dof / dt^2 (Mixture.F90:1224-1228);
    // a fresh start takes v0 = the box volume (Andersen_Initialize)
    const double an[8] = {pex, mass * dof / (s->dt * s->dt), v ? v[0] : box_volume(s->Lb),
                          v ? v[1] : 0.0, v ? v[2] : 0.0, v ? v[3] : 0.0, v ? v[4] : 0.0,
                          v ? v[5] : 0.0};
    if (!s->andersen) {
      double *box = nullptr, *anders = nullptr;
      if (hipMalloc((void **)&box, sizeof(double) * kBoxLen) != hipSuccess) return G6_EHIP;
      if (hipMalloc((void **)&anders, sizeof(double) * 8) != hipSuccess) {
        (void)hipFree(box);
        return G6_EHIP;
      }
      s->md.box = box;
      s->md.anders = anders;
      s->andersen = true;
    }
    G6_CHECK(hipStreamSynchronize(s->stream));
    G6_CHECK(hipMemcpy(s->md.box, bx, sizeof(bx), hipMemcpyHostToDevice));
    G6_CHECK(hipMemcpy(s->md.anders, an, sizeof(an), hipMemcpyHostToDevice));
    s->andersen_mode = mode;
    // rows of the new box
    int one[2] = {1, 1};
    G6_CHECK(hipMemcpy(s->md.flags + 2, one, sizeof(one), hipMemcpyHostToDevice));
  }
  drop_graphs(*s);  // the captured launches embed the step sequence
  return 0;
}

int g6_get_andersen(g6_handle h, double *st) {
  System *s = get(h);
  if (!s || !st) return G6_EARG;
  if (!s->andersen) {
    for (int q = 0; q < 8; q++) st[q] = 0.0;
    return 0;
  }
  double an[8];
  G6_CHECK(hipStreamSynchronize(s->stream));
  G6_CHECK(hipMemcpy(an, s->md.anders, sizeof(an), hipMemcpyDeviceToHost));
  st[0] = an[0];
  st[1] = an[1] * (s->dt * s->dt) / (6.0 * s->n);  // back to the @PCN4 mass
  for (int q = 2; q < 8; q++) st[q] = an[q];
  return 0;
}

int g6_get_box3(g6_handle h, double *edges) {
  System *s = get(h);
  if (!s || !edges) return G6_EARG;
  if (!s->andersen) {
    for (int a = 0; a < 3; a++) edges[a] = s->Lb[a];
    return 0;
  }
  G6_CHECK(hipStreamSynchronize(s->stream));
  G6_CHECK(hipMemcpy(edges, s->md.box, sizeof(double) * 3, hipMemcpyDeviceToHost));
  return 0;
}

int g6_get_box(g6_handle h, double *box_l) {
  if (!box_l) return G6_EARG;
  double e[3];
  const int rc = g6_get_box3(h, e);
  if (rc == 0) *box_l = e[0];
  return rc;
}

int g6_get_thermo(g6_handle h, double *thermo) {
  System *s = get(h);
  if (!s || !thermo) return G6_EARG;
  G6_CHECK(hipStreamSynchronize(s->stream));
  G6_CHECK(hipMemcpy(thermo, s->thermo, sizeof(double) * G6_THERMO_LEN, hipMemcpyDeviceToHost));
  return 0;
}

int g6_pairlist(g6_handle h, const double *com, int64_t cap, int64_t *npair0, int64_t *npair,
                int32_t *pair_i, int32_t *pair_j, double *ox, double *oy, double *oz) {
  System *s = get(h);
  if (!s || !com || !npair0 || !npair || cap < 0) return G6_EARG;
  if (s->skin != 0) return G6_EARG;  // the reference list is the skin-0 list
  return pairlist_once(*s, com, cap, npair0, npair, pair_i, pair_j, ox, oy, oz);
}

int g6_cells(g6_handle h, const double *com, int32_t *cell, int32_t *ndiv) {
  System *s = get(h);
  if (!s || !com || !cell || !ndiv) return G6_EARG;
  *ndiv = s->geo.nd[0];
  return cells_once(*s, com, cell);
}

int g6_cutoff(g6_handle h, int64_t npair, const int32_t *pi, const int32_t *pj,
              const double *com, const double *ox, const double *oy, const double *oz,
              double *eratio, double *fratio) {
  System *s = get(h);
  if (!s || npair < 0) return G6_EARG;
  return cutoff_once(*s, npair, pi, pj, com, ox, oy, oz, eratio, fratio);
}

int g6_force(g6_handle h, int64_t npair, const int32_t *pi, const int32_t *pj,
             const double *ox, const double *oy, const double *oz, const double *er,
             const double *fr, const double *x, const double *y, const double *z, double *fx,
             double *fy, double *fz, double *ep, double *vrmat) {
  System *s = get(h);
  if (!s || npair < 0 || !vrmat) return G6_EARG;
  return force_pairs(*s, npair, pi, pj, ox, oy, oz, er, fr, x, y, z, fx, fy, fz, ep, vrmat,
                     false);
}

int g6_pair_energy(g6_handle h, int64_t npair, const int32_t *pi, const int32_t *pj,
                   const double *ox, const double *oy, const double *oz, const double *er,
                   const double *x, const double *y, const double *z, double *ep) {
  System *s = get(h);
  if (!s || npair < 0) return G6_EARG;
  double v[9];
  return force_pairs(*s, npair, pi, pj, ox, oy, oz, er, nullptr, x, y, z, nullptr, nullptr,
                     nullptr, ep, v, true);
}

int g6_site_positions(g6_handle h, const double *com, const double *quat1, double *x,
                      double *y, double *z) {
  System *s = get(h);
  if (!s || !com || !quat1 || !x || !y || !z) return G6_EARG;
  return sites_once(*s, com, quat1, x, y, z);
}

int g6_collect(g6_handle h, const double *quat1, const double *fx, const double *fy,
               const double *fz, double *force, double *torque) {
  System *s = get(h);
  if (!s || !quat1 || !fx || !force || !torque) return G6_EARG;
  return collect_once(*s, quat1, fx, fy, fz, force, torque);
}

int g6_lattice_start(int n, double box_l, uint32_t seed, double *com, double *quat1) {
  if (n < 1 || !(box_l > 0) || !com || !quat1) return G6_EARG;
  int m = 1;
  while ((long long)m * m * m < n) m++;
  const long long sites = (long long)m * m * m;
  std::vector<int> perm((size_t)sites);
  for (long long k = 0; k < sites; k++) perm[k] = (int)k;
  MT mt(seed);
  // Fisher-Yates with the MT stream; the first n entries, sorted, are occupied
  for (long long k = sites - 1; k > 0; k--) {
    const long long r = (long long)(mt.uniform() * (double)(k + 1));
    const int t = perm[k];
    perm[k] = perm[r];
    perm[r] = t;
  }
  std::vector<int> occ(perm.begin(), perm.begin() + n);
  std::sort(occ.begin(), occ.end());
  const double a = box_l / m;
  for (int i = 0; i < n; i++) {
    const int k = occ[i];
    const int ix = k / (m * m), iy = (k / m) % m, iz = k % m;
    com[i] = (ix + 0.5) * a;
    com[n + i] = (iy + 0.5) * a;
    com[2 * n + i] = (iz + 0.5) * a;
    // Shoemake (1992): uniform on S^3 from three uniforms
    const double u1 = mt.uniform(), u2 = mt.uniform(), u3 = mt.uniform();
    quat1[i] = sqrt(1 - u1) * sin(2 * kPI * u2);
    quat1[n + i] = sqrt(1 - u1) * cos(2 * kPI * u2);
    quat1[2 * n + i] = sqrt(u1) * sin(2 * kPI * u3);
    quat1[3 * n + i] = sqrt(u1) * cos(2 * kPI * u3);
  }
  return 0;
}

}  // extern "C"
