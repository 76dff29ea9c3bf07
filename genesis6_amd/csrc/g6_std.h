// FP64 energy helpers shared by the Monte Carlo paths (g6_mc.hip, g6_npt.hip):
// StdPairPotential's site terms (StdInteraction.F90:633-762) with the
// CutOff_Dimmer2 switch (CutoffModule.F90:220-278), rigid_setmatrix and the
// site placement (RigidModule.F90:62-144).  The argument type A of the
// templates provides ns, bx/by/bz (body-frame sites) and rc, rin, c1.
#pragma once
#include <type_traits>

#include "g6_internal.h"

// FP64 energies in the reference's operation order: no FMA contraction
// (the pragma holds for the rest of every file that includes this one)
#pragma clang fp contract(off)

namespace g6 {

// The site pairs of StdPairPotential's term list for the MC models, in
// the order make_std_table produces them (checked at g6_mc_begin), so that
// the trial kernel indexes sites with constants.
template <int M>
struct StdPattern;
template <>
struct StdPattern<G6_TIP4P> {  // O-O LJ, then M,H,H x M,H,H Coulomb
  static constexpr int NT = 10, NS = 5;
  __device__ __host__ static constexpr int a(int u) { return u == 0 ? 0 : 1 + (u - 1) / 3; }
  __device__ __host__ static constexpr int b(int u) { return u == 0 ? 0 : 1 + (u - 1) % 3; }
  __device__ __host__ static constexpr int kind(int u) { return u == 0 ? 3 : 2; }
};
template <>
struct StdPattern<G6_TIP5P> {  // O-O LJ, then H,H,L,L x H,H,L,L Coulomb
  static constexpr int NT = 17, NS = 6;
  __device__ __host__ static constexpr int a(int u) { return u == 0 ? 0 : 1 + (u - 1) / 4; }
  __device__ __host__ static constexpr int b(int u) { return u == 0 ? 0 : 1 + (u - 1) % 4; }
  __device__ __host__ static constexpr int kind(int u) { return u == 0 ? 3 : 2; }
};
template <>
struct StdPattern<G6_SPCE> {  // O-O LJ + Coulomb, then the rest of O,H,H x O,H,H Coulomb
  static constexpr int NT = 9, NS = 4;
  __device__ __host__ static constexpr int a(int u) { return u / 3; }
  __device__ __host__ static constexpr int b(int u) { return u % 3; }
  __device__ __host__ static constexpr int kind(int u) { return u == 0 ? 1 : 2; }
};
// NvdE (O, H, H, M, L, L, CoM): every site pair with a charge product or an
// LJ eps product, site a of i outer and b of j inner
struct NvdeTerms {
  int a[30], b[30], k[30];
};
__device__ __host__ constexpr NvdeTerms nvde_terms() {
  NvdeTerms t{};
  const double q[6] = {0.0, 1.0, 1.0, 1.0, 1.0, 1.0};  // charged sites
  const double e[6] = {1.0, 1.0, 1.0, 0.0, 0.0, 0.0};  // LJ sites
  int u = 0;
  for (int a = 0; a < 6; a++)
    for (int b = 0; b < 6; b++) {
      const bool qq = q[a] * q[b] != 0.0, lj = e[a] * e[b] != 0.0;
      if (qq || lj) {
        t.a[u] = a;
        t.b[u] = b;
        t.k[u] = qq ? (lj ? 1 : 2) : 3;
        u++;
      }
    }
  return t;
}
template <>
struct StdPattern<G6_NVDE> {
  static constexpr int NT = 30, NS = 7;
  __device__ __host__ static constexpr int a(int u) { return nvde_terms().a[u]; }
  __device__ __host__ static constexpr int b(int u) { return nvde_terms().b[u]; }
  __device__ __host__ static constexpr int kind(int u) { return nvde_terms().k[u]; }
};

template <>
struct StdPattern<G6_OPLSMEOH> {  // Me, O, H x Me, O, H: LJ + Coulomb on Me/O pairs
  static constexpr int NT = 9, NS = 4;
  __device__ __host__ static constexpr int a(int u) { return u / 3; }
  __device__ __host__ static constexpr int b(int u) { return u % 3; }
  __device__ __host__ static constexpr int kind(int u) { return (u / 3 < 2 && u % 3 < 2) ? 1 : 2; }
};

// Calls f(std::integral_constant<int, M>) for the StdPairPotential term
// pattern M of a model: ST2's StdPairPotential terms (no Stillinger-Rahman
// switch, as mc2 evaluates them) have the TIP5P pattern.
template <class F>
inline void std_model(int model, F &&f) {
  if (model == G6_TIP4P) f(std::integral_constant<int, G6_TIP4P>());
  else if (model == G6_TIP5P || model == G6_ST2) f(std::integral_constant<int, G6_TIP5P>());
  else if (model == G6_SPCE) f(std::integral_constant<int, G6_SPCE>());
  else if (model == G6_OPLSMEOH) f(std::integral_constant<int, G6_OPLSMEOH>());
  else f(std::integral_constant<int, G6_NVDE>());
}

template <int M>
static bool std_table_matches(const StdTable &t) {
  if (t.nt != StdPattern<M>::NT) return false;
  for (int u = 0; u < t.nt; u++)
    if (t.t[u].a != StdPattern<M>::a(u) || t.t[u].b != StdPattern<M>::b(u) ||
        t.t[u].kind != StdPattern<M>::kind(u))
      return false;
  return true;
}

// CutOff_Dimmer2's S(r) on the COM distance squared (CutoffModule.F90:220-278)
template <class A>
__device__ inline double mc_switch(const A &a, double dd) {
  double hh = 0.0;
  if (dd < a.rc * a.rc) {
    if (dd > a.rin * a.rin) {
      const double dr = sqrt(dd), u = dr - a.rc, l = dr - a.rin, uu = u * u, ll = l * l;
      hh = a.c1 * uu * u * (10.0 * ll - 5.0 * u * l + uu);
    } else {
      hh = 1.0;
    }
  }
  return hh;
}

// StdPairPotential of one pair: site coordinates of i (xi..) and j (xj..),
// the image shift o subtracted from every site difference, times S.
__device__ inline double std_energy(const StdTable &st, const double *xi, const double *yi,
                                   const double *zi, const double *xj, const double *yj,
                                   const double *zj, double ox, double oy, double oz, double er) {
  double ep = 0.0;
  for (int u = 0; u < st.nt; u++) {
    const StdTerm s = st.t[u];
    double dx = xi[s.a] - xj[s.b], dy = yi[s.a] - yj[s.b], dz = zi[s.a] - zj[s.b];
    dx = dx - ox;
    dy = dy - oy;
    dz = dz - oz;
    const double dd = 1.0 / (dx * dx + dy * dy + dz * dz);
    double e0;
    if (s.kind == 1) {
      e0 = s.qq * sqrt(dd);
      const double d6 = dd * dd * dd, d12 = d6 * d6;
      e0 = e0 + s.aa * d12 + s.bb * d6;
    } else if (s.kind == 2) {
      e0 = s.qq * sqrt(dd);
    } else {
      const double d6 = dd * dd * dd, d12 = d6 * d6;
      e0 = s.aa * d12 + s.bb * d6;
    }
    ep = ep + e0;
  }
  return ep * er;
}

// std_energy with the term list of model M: j's sites from memory, the
// other molecule's sites in registers (constant indices)
template <int M>
__device__ inline double std_energy_t(const StdTable &st, const double *xj, const double *yj,
                                     const double *zj, const double (&xt)[kMaxSite],
                                     const double (&yt)[kMaxSite], const double (&zt)[kMaxSite],
                                     double ox, double oy, double oz, double er) {
  using P = StdPattern<M>;
  double ep = 0.0;
#pragma unroll
  for (int u = 0; u < P::NT; u++) {
    const StdTerm s = st.t[u];
    double dx = xj[P::a(u)] - xt[P::b(u)], dy = yj[P::a(u)] - yt[P::b(u)],
           dz = zj[P::a(u)] - zt[P::b(u)];
    dx = dx - ox;
    dy = dy - oy;
    dz = dz - oz;
    const double dd = 1.0 / (dx * dx + dy * dy + dz * dz);
    double e0;
    if (P::kind(u) == 1) {  // Coulomb + LJ
      e0 = s.qq * sqrt(dd);
      const double d6 = dd * dd * dd, d12 = d6 * d6;
      e0 = e0 + s.aa * d12 + s.bb * d6;
    } else if (P::kind(u) == 3) {  // LJ only
      const double d6 = dd * dd * dd, d12 = d6 * d6;
      e0 = s.aa * d12 + s.bb * d6;
    } else {
      e0 = s.qq * sqrt(dd);
    }
    ep = ep + e0;
  }
  return ep * er;
}

// std_energy_t with both molecules' sites in registers: a (outer, sites
// P::a) and b (inner, sites P::b), d = a - b - o
template <int M>
__device__ inline double std_energy_rr(const StdTable &st, const double (&xa)[kMaxSite],
                                      const double (&ya)[kMaxSite], const double (&za)[kMaxSite],
                                      const double (&xb)[kMaxSite], const double (&yb)[kMaxSite],
                                      const double (&zb)[kMaxSite], double ox, double oy,
                                      double oz, double er) {
  using P = StdPattern<M>;
  double ep = 0.0;
#pragma unroll
  for (int u = 0; u < P::NT; u++) {
    const StdTerm s = st.t[u];
    double dx = xa[P::a(u)] - xb[P::b(u)], dy = ya[P::a(u)] - yb[P::b(u)],
           dz = za[P::a(u)] - zb[P::b(u)];
    dx = dx - ox;
    dy = dy - oy;
    dz = dz - oz;
    const double dd = 1.0 / (dx * dx + dy * dy + dz * dz);
    double e0;
    if (P::kind(u) == 1) {  // Coulomb + LJ
      e0 = s.qq * sqrt(dd);
      const double d6 = dd * dd * dd, d12 = d6 * d6;
      e0 = e0 + s.aa * d12 + s.bb * d6;
    } else if (P::kind(u) == 3) {  // LJ only
      const double d6 = dd * dd * dd, d12 = d6 * d6;
      e0 = s.aa * d12 + s.bb * d6;
    } else {
      e0 = s.qq * sqrt(dd);
    }
    ep = ep + e0;
  }
  return ep * er;
}

// rigid_setmatrix (RigidModule.F90:62-83) and the site positions com + T b
// (rigid_setsiteposition0, RigidModule.F90:88-144)
template <class A>
__device__ inline void place_sites(const A &a, double qa, double qb, double qc, double qd,
                                   double cx, double cy, double cz, double *x, double *y,
                                   double *z) {
  const double t0 = (qa * qa + qb * qb - (qc * qc + qd * qd));
  const double t1 = -2.0 * (qa * qd + qb * qc);
  const double t2 = 2.0 * (qb * qd - qa * qc);
  const double t3 = 2.0 * (qa * qd - qb * qc);
  const double t4 = qa * qa + qc * qc - (qb * qb + qd * qd);
  const double t5 = -2.0 * (qa * qb + qc * qd);
  const double t6 = 2.0 * (qa * qc + qb * qd);
  const double t7 = 2.0 * (qa * qb - qc * qd);
  const double t8 = qa * qa + qd * qd - (qb * qb + qc * qc);
  for (int s = 0; s < a.ns; s++) {
    const double r0 = t0 * a.bx[s] + t1 * a.by[s] + t2 * a.bz[s];
    const double r1 = t3 * a.bx[s] + t4 * a.by[s] + t5 * a.bz[s];
    const double r2 = t6 * a.bx[s] + t7 * a.by[s] + t8 * a.bz[s];
    x[s] = cx + r0;
    y[s] = cy + r1;
    z[s] = cz + r2;
  }
}

// place_sites with a constant site count (the trial kernel's registers)
template <int NS, class A>
__device__ inline void place_sites_t(const A &a, double qa, double qb, double qc, double qd,
                                     double cx, double cy, double cz, double (&x)[kMaxSite],
                                     double (&y)[kMaxSite], double (&z)[kMaxSite]) {
  const double t0 = (qa * qa + qb * qb - (qc * qc + qd * qd));
  const double t1 = -2.0 * (qa * qd + qb * qc);
  const double t2 = 2.0 * (qb * qd - qa * qc);
  const double t3 = 2.0 * (qa * qd - qb * qc);
  const double t4 = qa * qa + qc * qc - (qb * qb + qd * qd);
  const double t5 = -2.0 * (qa * qb + qc * qd);
  const double t6 = 2.0 * (qa * qc + qb * qd);
  const double t7 = 2.0 * (qa * qb - qc * qd);
  const double t8 = qa * qa + qd * qd - (qb * qb + qc * qc);
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const double r0 = t0 * a.bx[s] + t1 * a.by[s] + t2 * a.bz[s];
    const double r1 = t3 * a.bx[s] + t4 * a.by[s] + t5 * a.bz[s];
    const double r2 = t6 * a.bx[s] + t7 * a.by[s] + t8 * a.bz[s];
    x[s] = cx + r0;
    y[s] = cy + r1;
    z[s] = cz + r2;
  }
}

}  // namespace g6
