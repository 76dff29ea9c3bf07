// Site-site interaction of one molecule pair, templated on the model and on
// the pair arithmetic type R (float: FP32-pair mode, double: FP64 mode).
//
// Restates, per pair k of the reference loops:
//   TIP4P  interaction_force_tip4p, TIP4PModule.F90:187-785 (9 Coulomb + O-O LJ,
//          scale by S, CoM term dh*d00*ep, ep1 = ep*S)
//   TIP5P  StdForce, StdInteraction.F90:201-576 (16 Coulomb + O-O LJ)
//   ST2    force_st2, ST2Module.F90:187-738 (16 Coulomb scaled by S*s12 with
//          the Stillinger-Rahman switch s12 on the COM distance, LJ by S)
//   SPC/E  StdForce (9 Coulomb over O, H, H + O-O LJ)
//   NvdE   StdForce (25 Coulomb over H, H, M, L, L + LJ on O-O, O-H, H-H)
//   OPLS methanol  StdForce (9 Coulomb over Me, O, H + LJ on Me-Me, Me-O,
//          O-Me, O-O)
// Cutoff ratios hh = S(r), dh = -S'(r)/r on the COM distance are inputs
// (cutoff_func, CutoffModule.F90:38-59; cutoff_sw below).
// Inputs are the COM separation d00 (shift already removed) and the site
// offsets from each COM; outputs are the per-site forces on i (fi) and the
// per-site forces that the reference subtracts from j (fj), and ep1.
#pragma once
#include "g6_internal.h"

namespace g6 {

template <int M>
struct Traits;
template <>
struct Traits<G6_TIP4P> {
  static constexpr int NS = 5, COM = 4, NC = 9, NLX = 0;
  static constexpr bool ST2SW = false;
  __device__ static constexpr int ci(int t) {
    constexpr int a[9] = {1, 1, 1, 2, 3, 2, 2, 3, 3};
    return a[t];
  }
  __device__ static constexpr int cj(int t) {
    constexpr int b[9] = {1, 2, 3, 1, 1, 2, 3, 2, 3};
    return b[t];
  }
};
template <>
struct Traits<G6_TIP5P> {
  static constexpr int NS = 6, COM = 5, NC = 16, NLX = 0;
  static constexpr bool ST2SW = false;
  __device__ static constexpr int ci(int t) { return 1 + t / 4; }
  __device__ static constexpr int cj(int t) { return 1 + t % 4; }
};
template <>
struct Traits<G6_ST2> {
  static constexpr int NS = 6, COM = 5, NC = 16, NLX = 0;
  static constexpr bool ST2SW = true;
  __device__ static constexpr int ci(int t) { return 1 + t / 4; }
  __device__ static constexpr int cj(int t) { return 1 + t % 4; }
};
template <>
struct Traits<G6_SPCE> {  // the O site carries both the charge and the LJ term
  static constexpr int NS = 4, COM = 3, NC = 9, NLX = 0;
  static constexpr bool ST2SW = false;
  __device__ static constexpr int ci(int t) { return t / 3; }
  __device__ static constexpr int cj(int t) { return t % 3; }
};
template <>
struct Traits<G6_NVDE> {  // O, H, H, M, L, L, CoM: LJ on O and H, charges on H, M, L
  static constexpr int NS = 7, COM = 6, NC = 25, NLX = 8;
  static constexpr bool ST2SW = false;
  __device__ static constexpr int ci(int t) { return 1 + t / 5; }
  __device__ static constexpr int cj(int t) { return 1 + t % 5; }
  // the further LJ pairs in the order of ModelTable::lx_a
  __device__ static constexpr int xa(int t) {
    constexpr int a[8] = {0, 0, 1, 2, 1, 1, 2, 2};
    return a[t];
  }
  __device__ static constexpr int xb(int t) {
    constexpr int b[8] = {1, 2, 0, 0, 1, 2, 1, 2};
    return b[t];
  }
};

template <>
struct Traits<G6_OPLSMEOH> {  // Me, O, H, CoM: LJ and charges on Me and O, a charge on H
  static constexpr int NS = 4, COM = 3, NC = 9, NLX = 3;
  static constexpr bool ST2SW = false;
  __device__ static constexpr int ci(int t) { return t / 3; }
  __device__ static constexpr int cj(int t) { return t % 3; }
  // the further LJ pairs in the order of ModelTable::lx_a: Me-O, O-Me, O-O
  __device__ static constexpr int xa(int t) {
    constexpr int a[3] = {0, 1, 1};
    return a[t];
  }
  __device__ static constexpr int xb(int t) {
    constexpr int b[3] = {1, 0, 1};
    return b[t];
  }
};

__device__ inline float g6_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline float g6_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Switching function S and dh = -S'/r between rin and rout.
template <typename R>
__device__ inline void cutoff_sw(const Geom &g, R dd, R &hh, R &dh) {
  hh = R(0);
  dh = R(0);
  if (dd < R(g.rout * g.rout)) {
    if (dd > R(g.rin * g.rin)) {
      const R dr = sqrt(dd);
      const R u = dr - R(g.rout), l = dr - R(g.rin);
      const R uu = u * u, ll = l * l;
      hh = R(g.c1) * uu * u * (R(10) * ll - R(5) * u * l + uu);
      dh = -(R(g.c2) * uu * ll / dr);
    } else {
      hh = R(1);
    }
  }
}

// 1/r^2 and 1/r of one site pair
template <typename R>
__device__ inline void inv_r(R r2, R &inv2, R &inv1);
template <>
__device__ inline void inv_r<double>(double r2, double &inv2, double &inv1) {
  inv2 = 1.0 / r2;
  inv1 = sqrt(inv2);
}
template <>
__device__ inline void inv_r<float>(float r2, float &inv2, float &inv1) {
  inv1 = g6_rsqrt(r2);
  inv2 = inv1 * inv1;
}

template <int M, typename R>
__device__ inline R pair_sites(const ModelTable &mt, R hh, R dh, R d0, R d1,
                               R d2, const R (&ri)[Traits<M>::NS][3],
                               const R (&rj)[Traits<M>::NS][3],
                               R (&fi)[Traits<M>::NS][3],
                               R (&fj)[Traits<M>::NS][3]) {
  using T = Traits<M>;
#pragma unroll
  for (int s = 0; s < T::NS; s++)
#pragma unroll
    for (int a = 0; a < 3; a++) fi[s][a] = fj[s][a] = R(0);
  R epc = R(0);
#pragma unroll
  for (int t = 0; t < T::NC; t++) {
    const int a = T::ci(t), b = T::cj(t);
    const R dx = d0 + (ri[a][0] - rj[b][0]);
    const R dy = d1 + (ri[a][1] - rj[b][1]);
    const R dz = d2 + (ri[a][2] - rj[b][2]);
    R inv2, inv1;
    inv_r<R>(dx * dx + dy * dy + dz * dz, inv2, inv1);
    const R e0 = R(mt.qq[t]) * inv1;
    const R f0 = e0 * inv2;
    epc += e0;
    fi[a][0] += dx * f0; fi[a][1] += dy * f0; fi[a][2] += dz * f0;
    fj[b][0] += dx * f0; fj[b][1] += dy * f0; fj[b][2] += dz * f0;
  }
  R eplj;
  {
    const R dx = d0 + (ri[0][0] - rj[0][0]);
    const R dy = d1 + (ri[0][1] - rj[0][1]);
    const R dz = d2 + (ri[0][2] - rj[0][2]);
    R inv2, inv1;
    inv_r<R>(dx * dx + dy * dy + dz * dz, inv2, inv1);
    const R d6 = inv2 * inv2 * inv2, d12 = d6 * d6;
    eplj = R(mt.lj_a) * d12 + R(mt.lj_b) * d6;
    const R f0 = inv2 * (R(mt.lj_a * 12.0) * d12 + R(mt.lj_b * 6.0) * d6);
    fi[0][0] += dx * f0; fi[0][1] += dy * f0; fi[0][2] += dz * f0;
    fj[0][0] += dx * f0; fj[0][1] += dy * f0; fj[0][2] += dz * f0;
  }
  if constexpr (T::NLX > 0) {
    static_assert(!T::ST2SW, "further LJ terms are scaled like the O-O term");
#pragma unroll
    for (int t = 0; t < T::NLX; t++) {
      const int a = T::xa(t), b = T::xb(t);
      const R dx = d0 + (ri[a][0] - rj[b][0]);
      const R dy = d1 + (ri[a][1] - rj[b][1]);
      const R dz = d2 + (ri[a][2] - rj[b][2]);
      R inv2, inv1;
      inv_r<R>(dx * dx + dy * dy + dz * dz, inv2, inv1);
      const R d6 = inv2 * inv2 * inv2, d12 = d6 * d6;
      eplj += R(mt.lx_a[t]) * d12 + R(mt.lx_b[t]) * d6;
      const R f0 = inv2 * (R(mt.lx_a[t] * 12.0) * d12 + R(mt.lx_b[t] * 6.0) * d6);
      fi[a][0] += dx * f0; fi[a][1] += dy * f0; fi[a][2] += dz * f0;
      fj[b][0] += dx * f0; fj[b][1] += dy * f0; fj[b][2] += dz * f0;
    }
  }
  R ep1;
  if constexpr (T::ST2SW) {
    const R dd = d0 * d0 + d1 * d1 + d2 * d2;
    R s12 = R(1), ds12 = R(0);
    const R rl = R(mt.rl), ru = R(mt.ru);
    if (dd < ru * ru) {
      if (dd < rl * rl) {
        s12 = R(0);
      } else {
        const R r12 = sqrt(dd);
        const R w3 = (ru - rl) * (ru - rl) * (ru - rl);
        s12 = (r12 - rl) * (r12 - rl) * (R(3) * ru - rl - R(2) * r12) / w3;
        ds12 = -(R(6) * (r12 - rl) * (ru - r12) / w3) / r12;
      }
    }
    const R hhc = hh * s12, dhc = dh * s12 + hh * ds12;
#pragma unroll
    for (int s = 0; s < T::NS; s++) {
      const R sc = (s == 0 || s == T::COM) ? hh : hhc;
#pragma unroll
      for (int a = 0; a < 3; a++) { fi[s][a] *= sc; fj[s][a] *= sc; }
    }
    const R cc = dh * eplj + dhc * epc;
    fi[T::COM][0] += cc * d0; fi[T::COM][1] += cc * d1; fi[T::COM][2] += cc * d2;
    fj[T::COM][0] += cc * d0; fj[T::COM][1] += cc * d1; fj[T::COM][2] += cc * d2;
    ep1 = epc * hhc + eplj * hh;
  } else {
#pragma unroll
    for (int s = 0; s < T::NS; s++)
#pragma unroll
      for (int a = 0; a < 3; a++) { fi[s][a] *= hh; fj[s][a] *= hh; }
    const R ep = epc + eplj;
    const R cc = dh * ep;
    fi[T::COM][0] += cc * d0; fi[T::COM][1] += cc * d1; fi[T::COM][2] += cc * d2;
    fj[T::COM][0] += cc * d0; fj[T::COM][1] += cc * d1; fj[T::COM][2] += cc * d2;
    ep1 = ep * hh;
  }
  return ep1;
}

// Force and torque on molecule i only (the MD kernel evaluates each pair
// from both sides) from the pair constants k: the same terms as pair_sites,
// with the switch factors applied to the summed site forces and torques
// instead of to each site force.  Site 0 carries the LJ term; the Coulomb
// terms sit on the sites from T::ci(0) up to the CoM site (from site 1 in
// TIP4P, TIP5P and ST2, from the O site in SPC/E).  dd = |d|^2 of the COM
// separation d.  Returns ep1.
template <int M, typename R>
__device__ inline R pair_on_i(const PairK<R> &k, R dd, R d0, R d1, R d2,
                              const R (&ri)[Traits<M>::NS][3],
                              const R (&rj)[Traits<M>::NS][3], R (&F)[3], R (&Tq)[3]) {
  using T = Traits<M>;
  constexpr int NQ = T::COM;  // sites 0..NQ-1 carry offsets
  // switch on the COM distance (cutoff_func, CutoffModule.F90:38-59); FP32
  // pairs take r and 1/r from one reciprocal square root
  R hh = R(0), dh = R(0);
  if (dd < k.rout2) {
    if (dd > k.rin2) {
      R dr, rinv;
      if constexpr (sizeof(R) == 4) {
        rinv = g6_rsqrt(dd);
        dr = dd * rinv;
      } else {
        dr = sqrt(dd);
        rinv = R(1) / dr;
      }
      const R u = dr - k.rout, l = dr - k.rin;
      const R uu = u * u, ll = l * l;
      hh = k.c1 * uu * u * (R(10) * ll - R(5) * u * l + uu);
      dh = -(k.c2 * uu * ll * rinv);
    } else {
      hh = R(1);
    }
  }
  // site positions of i relative to j's COM
  R pi[NQ][3];
#pragma unroll
  for (int s = 0; s < NQ; s++) {
    pi[s][0] = d0 + ri[s][0];
    pi[s][1] = d1 + ri[s][1];
    pi[s][2] = d2 + ri[s][2];
  }
  R fs[NQ][3];
#pragma unroll
  for (int s = 0; s < NQ; s++) fs[s][0] = fs[s][1] = fs[s][2] = R(0);
  R epc = R(0);
#pragma unroll
  for (int t = 0; t < T::NC; t++) {
    const int a = T::ci(t), b = T::cj(t);
    const R dx = pi[a][0] - rj[b][0];
    const R dy = pi[a][1] - rj[b][1];
    const R dz = pi[a][2] - rj[b][2];
    R inv2, inv1;
    inv_r<R>(dx * dx + dy * dy + dz * dz, inv2, inv1);
    const R e0 = k.qq[t] * inv1;
    const R f0 = e0 * inv2;
    epc += e0;
    fs[a][0] += dx * f0; fs[a][1] += dy * f0; fs[a][2] += dz * f0;
  }
  R eplj, fl[3];
  {
    const R dx = pi[0][0] - rj[0][0];
    const R dy = pi[0][1] - rj[0][1];
    const R dz = pi[0][2] - rj[0][2];
    R inv2, inv1;
    inv_r<R>(dx * dx + dy * dy + dz * dz, inv2, inv1);
    const R d6 = inv2 * inv2 * inv2, d12 = d6 * d6;
    eplj = k.lja * d12 + k.ljb * d6;
    const R f0 = inv2 * (k.lja12 * d12 + k.ljb6 * d6);
    fl[0] = dx * f0; fl[1] = dy * f0; fl[2] = dz * f0;
  }
  if constexpr (T::NLX > 0) {
    // further LJ terms: on site 0 with the O-O term, on the other sites with
    // the Coulomb sums (both scaled by S alone outside ST2)
    static_assert(!T::ST2SW, "further LJ terms are scaled like the O-O term");
#pragma unroll
    for (int t = 0; t < T::NLX; t++) {
      const int a = T::xa(t), b = T::xb(t);
      const R dx = pi[a][0] - rj[b][0];
      const R dy = pi[a][1] - rj[b][1];
      const R dz = pi[a][2] - rj[b][2];
      R inv2, inv1;
      inv_r<R>(dx * dx + dy * dy + dz * dz, inv2, inv1);
      const R d6 = inv2 * inv2 * inv2, d12 = d6 * d6;
      eplj += k.xa[t] * d12 + k.xb[t] * d6;
      const R f0 = inv2 * (k.xa12[t] * d12 + k.xb6[t] * d6);
      if (a == 0) {
        fl[0] += dx * f0; fl[1] += dy * f0; fl[2] += dz * f0;
      } else {
        fs[a][0] += dx * f0; fs[a][1] += dy * f0; fs[a][2] += dz * f0;
      }
    }
  }
  // Coulomb sums over the charged sites and the LJ site, unscaled
  R fc[3] = {R(0), R(0), R(0)}, tc[3] = {R(0), R(0), R(0)};
#pragma unroll
  for (int s = T::ci(0); s < NQ; s++) {
    fc[0] += fs[s][0]; fc[1] += fs[s][1]; fc[2] += fs[s][2];
    tc[0] += ri[s][1] * fs[s][2] - ri[s][2] * fs[s][1];
    tc[1] += ri[s][2] * fs[s][0] - ri[s][0] * fs[s][2];
    tc[2] += ri[s][0] * fs[s][1] - ri[s][1] * fs[s][0];
  }
  const R tl[3] = {ri[0][1] * fl[2] - ri[0][2] * fl[1],
                   ri[0][2] * fl[0] - ri[0][0] * fl[2],
                   ri[0][0] * fl[1] - ri[0][1] * fl[0]};
  R hc = hh, cc, ep1;
  if constexpr (T::ST2SW) {
    R s12 = R(1), ds12 = R(0);
    if (dd < k.ru2) {
      if (dd < k.rl2) {
        s12 = R(0);
      } else {
        R r12, rinv;
        if constexpr (sizeof(R) == 4) {
          rinv = g6_rsqrt(dd);
          r12 = dd * rinv;
        } else {
          r12 = sqrt(dd);
          rinv = R(1) / r12;
        }
        s12 = (r12 - k.rl) * (r12 - k.rl) * (R(3) * k.ru - k.rl - R(2) * r12) * k.w3inv;
        ds12 = -(R(6) * (r12 - k.rl) * (k.ru - r12) * k.w3inv) * rinv;
      }
    }
    hc = hh * s12;
    cc = dh * eplj + (dh * s12 + hh * ds12) * epc;
    ep1 = epc * hc + eplj * hh;
  } else {
    cc = dh * (epc + eplj);
    ep1 = (epc + eplj) * hh;
  }
  F[0] = hc * fc[0] + hh * fl[0] + cc * d0;
  F[1] = hc * fc[1] + hh * fl[1] + cc * d1;
  F[2] = hc * fc[2] + hh * fl[2] + cc * d2;
  Tq[0] = hc * tc[0] + hh * tl[0];
  Tq[1] = hc * tc[1] + hh * tl[1];
  Tq[2] = hc * tc[2] + hh * tl[2];
  return ep1;
}

// Energy only (PairPotential path): sum of site terms times S (and s12 for
// the ST2 Coulomb part, as force_st2's ep1).
template <int M, typename R>
__device__ inline R pair_energy(const ModelTable &mt, R hh, R d0, R d1,
                                R d2, const R (&ri)[Traits<M>::NS][3],
                                const R (&rj)[Traits<M>::NS][3]) {
  using T = Traits<M>;
  R epc = R(0);
#pragma unroll
  for (int t = 0; t < T::NC; t++) {
    const int a = T::ci(t), b = T::cj(t);
    const R dx = d0 + (ri[a][0] - rj[b][0]);
    const R dy = d1 + (ri[a][1] - rj[b][1]);
    const R dz = d2 + (ri[a][2] - rj[b][2]);
    R inv2, inv1;
    inv_r<R>(dx * dx + dy * dy + dz * dz, inv2, inv1);
    epc += R(mt.qq[t]) * inv1;
  }
  R eplj;
  {
    const R dx = d0 + (ri[0][0] - rj[0][0]);
    const R dy = d1 + (ri[0][1] - rj[0][1]);
    const R dz = d2 + (ri[0][2] - rj[0][2]);
    R inv2, inv1;
    inv_r<R>(dx * dx + dy * dy + dz * dz, inv2, inv1);
    const R d6 = inv2 * inv2 * inv2;
    eplj = R(mt.lj_a) * d6 * d6 + R(mt.lj_b) * d6;
  }
  if constexpr (T::NLX > 0) {
#pragma unroll
    for (int t = 0; t < T::NLX; t++) {
      const int a = T::xa(t), b = T::xb(t);
      const R dx = d0 + (ri[a][0] - rj[b][0]);
      const R dy = d1 + (ri[a][1] - rj[b][1]);
      const R dz = d2 + (ri[a][2] - rj[b][2]);
      R inv2, inv1;
      inv_r<R>(dx * dx + dy * dy + dz * dz, inv2, inv1);
      const R d6 = inv2 * inv2 * inv2;
      eplj += R(mt.lx_a[t]) * d6 * d6 + R(mt.lx_b[t]) * d6;
    }
  }
  if constexpr (T::ST2SW) {
    const R dd = d0 * d0 + d1 * d1 + d2 * d2;
    R s12 = R(1);
    const R rl = R(mt.rl), ru = R(mt.ru);
    if (dd < ru * ru) {
      if (dd < rl * rl) {
        s12 = R(0);
      } else {
        const R r12 = sqrt(dd);
        s12 = (r12 - rl) * (r12 - rl) * (R(3) * ru - rl - R(2) * r12) /
              ((ru - rl) * (ru - rl) * (ru - rl));
      }
    }
    return epc * hh * s12 + eplj * hh;
  } else {
    return (epc + eplj) * hh;
  }
}

}  // namespace g6
