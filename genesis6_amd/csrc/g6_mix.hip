// Mixtures of two rigid components (the reference's @NCMP 2, Mixture.F90:
// 475-499, 1833-2286): each component is a handle of its own (its molecules,
// model table, Gear state, rows and grid), and a mixture steps both with one
// Nose thermostat and adds the forces between them.
//
// Per step, on the first component's stream: each component's corrector and
// predictor with its binning (one k_md_pre launch each); the rows of each
// component (k_md_main, as a single component steps); then k_mix_cross, the
// pairs between the components, whose first launch carries the step close
// over both as one more block (the thermo record of the whole system: pair
// energies and virial of the three pair sets, the kinetic energies with each
// component's masses and moments, the temperature over all degrees of
// freedom, Nose_Correct and Nose_Predict, Mixture.F90:2236-2286).  Steps are
// captured in hipGraphs and replayed.  The second component reads
// the first's thermostat state, so both correctors take the same zeta.
//
// Pairs between components are StdForce's (StdInteraction.F90:63-630, the
// reference's force for every pair of a mixture without @FAST, Interaction2.
// F90:19-67): the site-pair terms of the two models' site parameters
// (Lorentz-Berthelot: sqrt(eps_a eps_b), (sig_a + sig_b)/2, q_a q_b), times
// the switch S on the COM distance, with the switch's derivative on the COMs.
// k_mix_cross takes every molecule of one component against every molecule
// of the other (a 16-lane row, a wave or a block per molecule, its lanes over
// the other component's molecules in index order; both directions, every sum
// in a fixed order, so the step is deterministic), and adds each molecule's
// force and lab torque to its component's force array.
#include <math.h>
#include <stdlib.h>

#include <initializer_list>
#include <mutex>
#include <vector>

#include "g6_internal.h"
#include "g6_md_common.h"
#include "g6_std.h"

#pragma clang fp contract(off)

namespace g6 {

// site terms between the components: site a of the row component, site b of
// the other (qq, aa, bb as StdInteraction.F90:201-221 forms them)
struct CrossArgs {
  int ni, nj, nsi, nsj;    // molecules and sites (CoM site last) of each side
  const double *reci;      // records of the row component (COM, site offsets)
  const double *recj;
  const double *comj;      // the other component's COMs, [3][nj] (the records' values)
  int rs;                  // record stride (doubles): kRecD FP64 offsets, kRecF FP32
  double *frc;             // the row component's force and lab torque [6][ni]
  double *part;            // [nblk][kBP] pair partials of the launch
  double L[3], invL[3];
  double rc2, rin, rout, c1, c2;
  StdTable t;
  signed char term[kMaxSite][kMaxSite];  // index into t of site pair (a, b), -1 none
};

constexpr int kMixThreads = 256;  // four waves, one molecule each
constexpr int kMixCap = 512;      // compacted pairs a wave holds before the site terms

// site offset q (site * 3 + axis) of a record: FP64 records hold them as
// doubles after the COM, FP32 records as floats (the pair precision's
// rounding of the same lab vectors)
__device__ inline double site_off(const double *rec, int rs, int q) {
  return rs == kRecF ? (double)reinterpret_cast<const float *>(rec + 4)[q] : rec[4 + q];
}

// The step close of a mixture (the single-component step close for two
// components and the pairs between them).  mode: 1 thermo record and
// Nose_Correct, 2 Nose_Predict, 4 clear the cells and rebuild flag of parity
// 1 - parity of both components, 8 of both parities.  The thermo record of
// the first component takes the whole system: energy and virial summed over
// the pair sets (first component, second, between) in that order, as
// Mixture.F90:2085-2090 sums ep(k) and virialp(k); kinetic energies and
// tensors summed over the components, each with its own mass and moments
// (Mixture.F90:2180-2212); the temperature over 6 (nA + nB) degrees of
// freedom.  parts: ep of each set and ek of each component (internal units).
// What the close of each component reads and writes (small enough to ride
// in the cross kernel's arguments beside its site table).
struct MixSide {
  double mass, Ixx, Iyy, Izz;
  const double *blockek;
  int *cellres, *flags;
  int cell_cap;
  double half_skin2;
};
struct MixClose {
  MixSide a, b;
  double *nose, *thermo;  // the first component's
  int na, nb, nblk_pre_a, nblk_pre_b, nblk_main_a, nblk_main_b, nblk_x, ncell_a, ncell_b;
  const double *bpart_a, *bpart_b, *bpart_x;
  double *parts;
  double dt, box_vol;
  int nose_active;
};

constexpr int kMixRed = 36 + 2 * kEkLen;

__device__ inline void mix_close_body(const MixClose &c, int mode, int parity) {
  __shared__ double s_red[4][kMixRed];
  const int wid = threadIdx.x >> 6;
  {
    // pair partials of each set, then kinetic partials of each component,
    // per thread in block order, then over the wave and the four waves
    double v[kMixRed];
#pragma unroll
    for (int q = 0; q < kMixRed; q++) v[q] = 0.0;
    if (mode & 1) {
      for (int k = threadIdx.x; k < c.nblk_main_a; k += 256)
#pragma unroll
        for (int q = 0; q < 12; q++) v[q] += c.bpart_a[(size_t)k * kBP + q];
      for (int k = threadIdx.x; k < c.nblk_main_b; k += 256)
#pragma unroll
        for (int q = 0; q < 12; q++) v[12 + q] += c.bpart_b[(size_t)k * kBP + q];
      for (int k = threadIdx.x; k < c.nblk_x; k += 256)
#pragma unroll
        for (int q = 0; q < 12; q++) v[24 + q] += c.bpart_x[(size_t)k * kBP + q];
      for (int k = threadIdx.x; k < c.nblk_pre_a; k += 256)
#pragma unroll
        for (int q = 0; q < kEkLen; q++) v[36 + q] += c.a.blockek[(size_t)k * kEkLen + q];
      for (int k = threadIdx.x; k < c.nblk_pre_b; k += 256)
#pragma unroll
        for (int q = 0; q < kEkLen; q++) v[36 + kEkLen + q] += c.b.blockek[(size_t)k * kEkLen + q];
    }
#pragma unroll
    for (int q = 0; q < kMixRed; q++) v[q] = wsum63(v[q]);
    if ((threadIdx.x & 63) == 63)
#pragma unroll
      for (int q = 0; q < kMixRed; q++) s_red[wid][q] = v[q];
  }
  if (mode & 12) {
    const int p0 = (mode & 8) ? 0 : 1 - parity, p1 = (mode & 8) ? 1 : 1 - parity;
    for (int p = p0; p <= p1; p++)
      for (int u = 0; u < 2; u++) {
        const MixSide &m = u ? c.b : c.a;
        const int ncell = u ? c.ncell_b : c.ncell_a;
        int *cr = m.cellres + (size_t)p * ncell * m.cell_cap;
        for (int q = threadIdx.x; q < ncell; q += 256) cr[(size_t)q * m.cell_cap] = 0;
        if (threadIdx.x == 0 && m.half_skin2 > 0.0) m.flags[2 + p] = 0;
      }
  }
  __syncthreads();
  if (threadIdx.x < kMixRed) {
    const int q = threadIdx.x;
    s_red[0][q] = s_red[0][q] + s_red[1][q] + s_red[2][q] + s_red[3][q];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double *ra = s_red[0], *rb = s_red[0] + 12, *rx = s_red[0] + 24;
  const double *ka = s_red[0] + 36, *kb = s_red[0] + 36 + kEkLen;
  double *z = c.nose, *th = c.thermo;
  const double dt = c.dt;
  if (mode & 1) {
    th[0] = (ra[0] + rb[0]) + rx[0];
    for (int q = 0; q < 9; q++) th[4 + q] = (ra[1 + q] + rb[1 + q]) + rx[1 + q];
    th[14] = (ra[10] + rb[10]) + rx[10];
    const double nl = ra[11] + rb[11];
    if (nl > 0) th[15] = nl;
    // Rigid_Ek of each component (RigidModule.F90:576-599)
    const double kd2 = 0.5 * (1.0 / (dt * dt));
    const double ekta = ka[0] * c.a.mass * kd2, ektb = kb[0] * c.b.mass * kd2;
    const double ekra = (ka[1] * c.a.Ixx + ka[2] * c.a.Iyy + ka[3] * c.a.Izz) * 0.5;
    const double ekrb = (kb[1] * c.b.Ixx + kb[2] * c.b.Iyy + kb[3] * c.b.Izz) * 0.5;
    const double ekt = ekta + ektb, ekr = ekra + ekrb;
    const double dof = 6.0 * (c.na + c.nb);
    const double temp = 10.0 * 2.0 * (ekt + ekr) / (G6_RGAS_MD * dof);
    double h = (th[0] + (ekt + ekr)) * 10.0;
    double nose_ep = 0.0, nose_ek = 0.0;
    if (c.nose_active) {
      const double zetaf = (temp - z[0]) / z[1] - z[2] * z[3];
      const double dz2 = zetaf - z[4];
      z[3] = z[3] + kG4_0 * dz2;
      z[4] = zetaf;
      z[5] = z[5] + kG4_2 * dz2;
      z[6] = z[6] + kG4_3 * dz2;
      z[7] = z[7] + kG4_4 * dz2;
      z[8] = z[8] + z[3];
      nose_ek = (z[3] * z[3] * z[1] * dt * dof * 0.5) * G6_RGAS_MD;
      nose_ep = (dof * z[0] * (z[8] - 0.5 * z[3]) * dt) * G6_RGAS_MD;
      h = h + nose_ep + nose_ek;
    }
    th[1] = ekt;
    th[2] = ekr;
    th[3] = temp;
    th[13] = h;
    // kinetic tensors (RigidModule.F90:617), summed over the components,
    // and the pressure (Mixture.F90:2206-2220)
    const double kd = 1.0 / (dt * dt);
    const int idx[9] = {4, 5, 6, 5, 7, 8, 6, 8, 9};
    double kt[9];
    for (int q = 0; q < 9; q++) kt[q] = ka[idx[q]] * c.a.mass * kd + kb[idx[q]] * c.b.mass * kd;
    for (int q = 0; q < 9; q++) th[16 + q] = kt[q];
    const double vp = (kt[0] + th[4]) + (kt[4] + th[8]) + (kt[8] + th[12]);
    th[25] = (vp / 3.0) * 10.0 * (1.0 / 101326.0) /
             (c.box_vol * 6.0225e23 * (1e-10 * 1e-10 * 1e-10));
    th[26] = c.box_vol;
    th[27] = nose_ep;
    th[28] = nose_ek;
    th[29] = c.nose_active ? z[0] : 0.0;
    th[30] = 0.0;
    th[31] = 0.0;
    c.parts[0] = ra[0];
    c.parts[1] = rb[0];
    c.parts[2] = rx[0];
    c.parts[3] = ekta + ekra;
    c.parts[4] = ektb + ekrb;
  }
  if ((mode & 2) && c.nose_active) {
    z[3] = z[3] + z[4] + z[5] + z[6] + z[7];
    z[4] = z[4] + 2.0 * z[5] + 3.0 * z[6] + 4.0 * z[7];
    z[5] = z[5] + 3.0 * z[6] + 6.0 * z[7];
    z[6] = z[6] + 4.0 * z[7];
  }
}

__global__ __launch_bounds__(256) void k_mix_close(MixClose c, int mode, int parity) {
  mix_close_body(c, mode, parity);
}

// For each molecule i of the row component (a group of lanes: see LPM
// below), the other component's molecules are taken in chunks in index
// order: the COM test dd < rc^2 (minimum image per axis) on all of them, the
// indices of the pairs that pass appended to the group's LDS row in index
// order; whenever a row is nearly full, and at the end, the pairs are
// evaluated densely over the group's lanes (the COM vector again, S and dS
// on the COM distance, StdForce's site terms), so the site terms run once
// per group-width of pairs within rc rather than once per group-width of
// candidates.  Force, lab torque, energy, virial and pair count are summed
// per lane, then over the group in a fixed order; i's force and torque are
// added to its force array, and the block's energy, virial and count (its
// groups in order, halved: every pair is met from both sides) go to its
// partials.

// sum over the LPM lanes of a molecule's group in a fixed order (the whole
// wave for 64; a 16-lane row for 16, every lane of the row receiving it)
template <int LPM>
__device__ inline double group_sum(double v) {
  if constexpr (LPM == 64) return wsum63(v);
  else return rowsum16(v);
}

// LPM: lanes per molecule.  16: four molecules per wave (a 16-lane row
// each); 64: a wave per molecule; 256: a block per molecule, each wave
// taking a quarter of the candidates (contiguous index ranges) and the four
// wave totals summed in wave order.
template <int LPM>
__global__ __launch_bounds__(kMixThreads) void k_mix_cross(CrossArgs x, MixClose c, int close_mode,
                                                           int close_parity) {
  static_assert(LPM == 16 || LPM == 64 || LPM == 256, "16, 64 or 256 lanes per molecule");
  static_assert(kMixThreads == 256, "the step close runs on 256 threads");
  // the last block of a launch with close_mode != 0 closes the step (it
  // touches nothing the cross blocks read or write)
  if (close_mode && blockIdx.x == gridDim.x - 1) {
    mix_close_body(c, close_mode, close_parity);
    return;
  }
  constexpr bool SPLIT = LPM == 256;
  constexpr int GL = SPLIT ? 64 : LPM;  // lanes of a molecule within a wave
  constexpr int MPW = 64 / GL;          // molecules per wave
  constexpr int CAP = kMixCap / MPW;    // LDS row entries per group
  constexpr int CH = 4 * GL;            // candidates per pass per group
  __shared__ double s_v[kMixThreads / 64 * MPW][12 + 6];
  __shared__ int s_j[kMixThreads / 64][kMixCap];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int grp = lane / GL, sub = lane % GL;
  const int slot = wid * MPW + grp;
  const int i = SPLIT ? (int)blockIdx.x : (int)blockIdx.x * (kMixThreads / 64 * MPW) + slot;
  const bool live = i < x.ni;
  // this group's candidates: all of the other component, or the wave's
  // quarter of it
  const int jb = SPLIT ? (int)((long long)x.nj * wid / 4) : 0;
  const int je = SPLIT ? (int)((long long)x.nj * (wid + 1) / 4) : x.nj;
  double acc[12];
#pragma unroll
  for (int q = 0; q < 12; q++) acc[q] = 0.0;
  double F[3] = {0, 0, 0}, Tq[3] = {0, 0, 0};
  if (__ballot(live) != 0) {  // uniform over the wave; groups past ni find no pairs
    // site offsets in registers: every site index below is a compile-time
    // constant (unrolled loops over kMaxSite, the term table uniform); the
    // CoM site and the sites past a model's count are zero
    const double *ri = x.reci + (size_t)(live ? i : 0) * x.rs;
    double si[kMaxSite][3];
#pragma unroll
    for (int s = 0; s < kMaxSite; s++)
#pragma unroll
      for (int a = 0; a < 3; a++) si[s][a] = s + 1 < x.nsi ? site_off(ri, x.rs, s * 3 + a) : 0.0;
    const double xi = ri[0], yi = ri[1], zi = ri[2];
    int *cj = s_j[wid] + grp * CAP;
    int np = 0;  // this group's entries (equal over its lanes)
    for (int j0 = jb; j0 < je; j0 += CH) {
      // pass 1: the COM test on the chunk's candidates, four per lane with
      // their loads in flight together; the pairs within rc are appended in
      // index order to this wave's LDS rows
      {
        double xj[4], yj[4], zj[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int jr = j0 + GL * u + sub, j = jr < je ? jr : 0;
          xj[u] = x.comj[j];
          yj[u] = x.comj[x.nj + j];
          zj[u] = x.comj[2 * x.nj + j];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int j = j0 + GL * u + sub;
          double d0 = xi - xj[u], d1 = yi - yj[u], d2 = zi - zj[u];
          d0 = d0 - round(d0 * x.invL[0]) * x.L[0];
          d1 = d1 - round(d1 * x.invL[1]) * x.L[1];
          d2 = d2 - round(d2 * x.invL[2]) * x.L[2];
          const bool pass = live && j < je && d0 * d0 + d1 * d1 + d2 * d2 < x.rc2;
          unsigned long long mask = __ballot(pass);
          if constexpr (GL < 64) mask = (mask >> (grp * GL)) & ((1ull << GL) - 1ull);
          if (pass) cj[np + __popcll(mask & ((1ull << sub) - 1ull))] = j;
          np += __popcll(mask);
        }
      }
      wsync();
      // pass 2, once a row cannot take another chunk or the candidates are
      // done (a decision of the whole wave): the pairs held, densely over
      // each group's lanes
      if (__ballot(np > CAP - CH) == 0 && j0 + CH < je) continue;
      for (int k = sub; k < np; k += GL) {
        const int j = cj[k];
        const double *rj = x.recj + (size_t)j * x.rs;
        // the COM vector again, by pass 1's expression (the same bits)
        double d[3] = {xi - x.comj[j], yi - x.comj[x.nj + j], zi - x.comj[2 * x.nj + j]};
        for (int a = 0; a < 3; a++) d[a] = d[a] - round(d[a] * x.invL[a]) * x.L[a];
        const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        // cutoff_func (CutoffModule.F90:38-59) on the COM distance
        double hh = 1.0, dh = 0.0;
        if (dd > x.rin * x.rin) {
          const double dr = sqrt(dd), u = dr - x.rout, l = dr - x.rin, uu = u * u, ll = l * l;
          hh = x.c1 * uu * u * (10.0 * ll - 5.0 * u * l + uu);
          dh = -(x.c2 * uu * ll / dr);
        }
        double sj[kMaxSite][3];
#pragma unroll
        for (int s = 0; s < kMaxSite; s++)
#pragma unroll
          for (int a = 0; a < 3; a++) sj[s][a] = s + 1 < x.nsj ? site_off(rj, x.rs, s * 3 + a) : 0.0;
        double fs[kMaxSite][3];
#pragma unroll
        for (int s = 0; s < kMaxSite; s++) fs[s][0] = fs[s][1] = fs[s][2] = 0.0;
        double ep = 0.0;
        // the terms in table order (site a of i outer, b of j inner)
#pragma unroll
        for (int a = 0; a < kMaxSite; a++)
#pragma unroll
        for (int b = 0; b < kMaxSite; b++) {
          const int u = x.term[a][b];
          if (u < 0) continue;
          const StdTerm &tm = x.t.t[u];
          const double dx = (d[0] + si[a][0]) - sj[b][0];
          const double dy = (d[1] + si[a][1]) - sj[b][1];
          const double dz = (d[2] + si[a][2]) - sj[b][2];
          const double inv2 = 1.0 / (dx * dx + dy * dy + dz * dz);
          double e0 = 0.0, f0 = 0.0;
          if (tm.kind != 3) {
            const double ec = tm.qq * sqrt(inv2);
            e0 = ec;
            f0 = ec * inv2;
          }
          if (tm.kind != 2) {
            const double d6 = inv2 * inv2 * inv2, d12 = d6 * d6;
            e0 = e0 + tm.aa * d12 + tm.bb * d6;
            f0 = f0 + inv2 * (12.0 * tm.aa * d12 + 6.0 * tm.bb * d6);
          }
          ep = ep + e0;
          fs[a][0] += dx * f0;
          fs[a][1] += dy * f0;
          fs[a][2] += dz * f0;
        }
        double fc[3] = {0, 0, 0}, tc[3] = {0, 0, 0};
#pragma unroll
        for (int s = 0; s < kMaxSite; s++) {
          fc[0] += fs[s][0];
          fc[1] += fs[s][1];
          fc[2] += fs[s][2];
          tc[0] += si[s][1] * fs[s][2] - si[s][2] * fs[s][1];
          tc[1] += si[s][2] * fs[s][0] - si[s][0] * fs[s][2];
          tc[2] += si[s][0] * fs[s][1] - si[s][1] * fs[s][0];
        }
        const double cc = dh * ep;
        double Fp[3];
        for (int a = 0; a < 3; a++) {
          Fp[a] = hh * fc[a] + cc * d[a];
          F[a] += Fp[a];
          Tq[a] += hh * tc[a];
        }
        acc[0] += ep * hh;
        for (int a = 0; a < 3; a++)
          for (int c = 0; c < 3; c++) acc[1 + 3 * a + c] = fma(d[a], Fp[c], acc[1 + 3 * a + c]);
        acc[10] += 1.0;
      }
      wsync();  // the next chunk overwrites the rows
      np = 0;
    }
  }
#pragma unroll
  for (int q = 0; q < 12; q++) acc[q] = group_sum<GL>(acc[q]);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    F[a] = group_sum<GL>(F[a]);
    Tq[a] = group_sum<GL>(Tq[a]);
  }
  if (sub == GL - 1) {
    if (live && !SPLIT) {
      const int ni = x.ni;
      for (int a = 0; a < 3; a++) {
        x.frc[a * ni + i] = x.frc[a * ni + i] + F[a];
        x.frc[(3 + a) * ni + i] = x.frc[(3 + a) * ni + i] + Tq[a];
      }
    }
#pragma unroll
    for (int q = 0; q < 12; q++) s_v[slot][q] = acc[q];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      s_v[slot][12 + a] = F[a];
      s_v[slot][15 + a] = Tq[a];
    }
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    double v = 0.0;
    for (int w = 0; w < kMixThreads / 64 * MPW; w++) v += s_v[w][threadIdx.x];
    x.part[(size_t)blockIdx.x * kBP + threadIdx.x] = 0.5 * v;
  } else if (SPLIT && threadIdx.x < 18 && live) {
    // the molecule's force and torque: the four waves' totals in order
    const int q = threadIdx.x - 12, ni = x.ni;
    double v = 0.0;
    for (int w = 0; w < 4; w++) v += s_v[w][12 + q];
    x.frc[q * ni + i] = x.frc[q * ni + i] + v;
  }
}


struct Mix {
  g6_handle ha = 0, hb = 0;
  System *a = nullptr, *b = nullptr;
  StdTable ab{}, ba{};       // site terms: a's sites with b's, b's with a's
  double *part = nullptr;    // [2][nblk_a + nblk_b][kBP]: cross partials by parity
  double *parts = nullptr;   // [8]: ep of each pair set, ek of each component
  int nblk_a = 0, nblk_b = 0;
  int lpm_a = 64, lpm_b = 64;  // lanes per molecule of each direction
  // captured step sequences by step count (four at most), valid while
  // neither component dropped its own graphs (graph_gen) since the capture
  hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
  int graph_k[4] = {0, 0, 0, 0};
  int graph_next = 0, gen_a = -1, gen_b = -1;
};

static void mix_drop_graphs(Mix &m) {
  for (int q = 0; q < 4; q++) {
    if (m.graph[q]) (void)hipGraphExecDestroy(m.graph[q]);
    m.graph[q] = nullptr;
    m.graph_k[q] = 0;
  }
}

static std::mutex g_mxmu;
static std::vector<Mix *> g_mix;

static Mix *mix_of(int64_t h) {
  std::lock_guard<std::mutex> lk(g_mxmu);
  if (h <= 0 || h > (int64_t)g_mix.size()) return nullptr;
  return g_mix[h - 1];
}

static CrossArgs cross_args(const System &r, const System &o, const StdTable &t, double *part) {
  CrossArgs x;
  x.ni = r.n;
  x.nj = o.n;
  x.nsi = r.mt.nsite;
  x.nsj = o.mt.nsite;
  x.reci = r.md.rec;
  x.recj = o.md.rec;
  x.comj = o.md.com;
  x.rs = r.md.rs;
  x.frc = r.md.frc;
  x.part = part;
  for (int a = 0; a < 3; a++) {
    x.L[a] = r.Lb[a];
    x.invL[a] = 1.0 / r.Lb[a];
  }
  x.rc2 = r.geo_md.rc2;
  x.rin = r.geo_md.rin;
  x.rout = r.geo_md.rout;
  x.c1 = r.geo_md.c1;
  x.c2 = r.geo_md.c2;
  x.t = t;
  for (int a = 0; a < kMaxSite; a++)
    for (int b = 0; b < kMaxSite; b++) x.term[a][b] = -1;
  for (int u = 0; u < t.nt; u++) x.term[t.t[u].a][t.t[u].b] = (signed char)u;
  return x;
}

static MixSide mix_side(const System &s) {
  MixSide d;
  d.mass = s.mt.mass;
  d.Ixx = s.mt.Ixx;
  d.Iyy = s.mt.Iyy;
  d.Izz = s.mt.Izz;
  d.blockek = s.md.blockek;
  d.cellres = s.md.cellres;
  d.flags = s.md.flags;
  d.cell_cap = s.md.cell_cap;
  d.half_skin2 = s.md.half_skin2;
  return d;
}

static MixClose mix_close_args(const Mix &m, int src_parity) {
  const System &a = *m.a, &b = *m.b;
  MixClose c;
  c.a = mix_side(a);
  c.b = mix_side(b);
  c.nose = a.md.nose;
  c.thermo = a.md.thermo;
  c.na = a.n;
  c.nb = b.n;
  c.nblk_pre_a = pre_waves(a.n);
  c.nblk_pre_b = pre_waves(b.n);
  c.nblk_main_a = a.nblk_md;
  c.nblk_main_b = b.nblk_md;
  c.nblk_x = m.nblk_a + m.nblk_b;
  c.ncell_a = a.geo_md.ncell;
  c.ncell_b = b.geo_md.ncell;
  c.bpart_a = a.md.blockpart + (size_t)src_parity * a.nblk_md * kBP;
  c.bpart_b = b.md.blockpart + (size_t)src_parity * b.nblk_md * kBP;
  c.bpart_x = m.part + (size_t)src_parity * (m.nblk_a + m.nblk_b) * kBP;
  c.parts = m.parts;
  c.dt = a.dt;
  c.box_vol = box_volume(a.Lb);
  c.nose_active = a.nose_active ? 1 : 0;
  return c;
}

static void mix_close(Mix &m, int mode, int parity, int src_parity) {
  k_mix_close<<<1, 256, 0, m.a->stream>>>(mix_close_args(m, src_parity), mode, parity);
}

// the cross pairs of step parity p; the first launch carries one more block
// that closes the step (mode close_mode, src parity 1 - p) when close_mode
// is set
static void mix_cross(Mix &m, int parity, int close_mode) {
  double *p = m.part + (size_t)parity * (m.nblk_a + m.nblk_b) * kBP;
  const CrossArgs xa = cross_args(*m.a, *m.b, m.ab, p);
  const CrossArgs xb = cross_args(*m.b, *m.a, m.ba, p + (size_t)m.nblk_a * kBP);
  const MixClose c = mix_close_args(m, 1 - parity);
  auto launch = [&](int lpm, int nblk, const CrossArgs &x, int mode) {
    const int g = nblk + (mode ? 1 : 0);
    if (lpm == 16) k_mix_cross<16><<<g, kMixThreads, 0, m.a->stream>>>(x, c, mode, parity);
    else if (lpm == 256) k_mix_cross<256><<<g, kMixThreads, 0, m.a->stream>>>(x, c, mode, parity);
    else k_mix_cross<64><<<g, kMixThreads, 0, m.a->stream>>>(x, c, mode, parity);
  };
  launch(m.lpm_a, m.nblk_a, xa, close_mode);
  launch(m.lpm_b, m.nblk_b, xb, 0);
}

// k mixture steps, every launch on the first component's stream (the
// phases of the single-component step: corrector and predictor, close,
// forces)
static void mix_steps(Mix &m, int k) {
  System &a = *m.a, &b = *m.b;
  for (int q = 0; q < k; q++) {
    const int p = q & 1;
    // correctors (from the second step on) and predictors in one launch
    // per component, binned into the cells of parity p (the predictors do
    // not read the thermostat, which the close below updates)
    md_launch_pre(a, q > 0, true, true, p);
    md_launch_pre(b, q > 0, true, true, p);
    md_launch_main(a, nullptr, p, 0);
    md_launch_main(b, nullptr, p, 0);
    // the close rides in the first cross launch: it reads the correctors'
    // kinetic partials and the previous step's pair partials, updates the
    // thermostat (read by the next correctors) and clears the cells of
    // parity 1 - p; the rows and cross pairs of this step use none of these
    mix_cross(m, p, (q > 0 ? 1 : 0) | 2 | 4);
  }
  md_launch_pre(a, true, false, false, 0);
  md_launch_pre(b, true, false, false, 0);
  mix_close(m, 1 | 8, 0, (k - 1) & 1);
}

// both components still alive (neither handle finalized since the create)
static bool mix_live(const Mix &m) {
  return system_of(m.ha) == m.a && system_of(m.hb) == m.b;
}

static int mix_flags(System &s) {
  int fl = 0;
  G6_CHECK(hipMemcpyAsync(&fl, s.md.flags, sizeof(int), hipMemcpyDeviceToHost, s.stream));
  G6_CHECK(hipStreamSynchronize(s.stream));
  s.last_flag = fl;
  if (fl == 2) return G6_EARG;
  return fl ? G6_ECAPACITY : 0;
}

}  // namespace g6

using namespace g6;

extern "C" {

int g6_mixture_create(g6_mixture *hm, g6_handle ha, g6_handle hb) {
  if (!hm) return G6_EARG;
  System *a = system_of(ha), *b = system_of(hb);
  if (!a || !b || a == b) return G6_EARG;
  // one box, cutoff, time step, pair-list skin and pair precision; a fixed
  // box; the rows of both directions (a component on the cell tiles moves to
  // them, as a ladder's systems do)
  for (int k = 0; k < 3; k++)
    if (a->Lb[k] != b->Lb[k]) return G6_EARG;
  if (a->rc != b->rc || a->taper != b->taper || a->dt != b->dt || a->skin != b->skin ||
      a->prec != b->prec || a->andersen || b->andersen)
    return G6_EARG;
  {
    std::lock_guard<std::mutex> lk(g_mxmu);
    for (Mix *o : g_mix)
      if (o && (o->a == a || o->b == a || o->a == b || o->b == b)) return G6_EARG;
  }
  Mix *m = new Mix();
  m->ha = ha;
  m->hb = hb;
  m->a = a;
  m->b = b;
  make_std_table(a->mt, b->mt, &m->ab);
  make_std_table(b->mt, a->mt, &m->ba);
  // a 16-lane row per molecule where a molecule has few partners of the
  // other component within rc (expected count nj (4/3) pi rc^3 / V below
  // 24), else a wave
  const double vol = box_volume(a->Lb), sph = 4.0 / 3.0 * 3.14159265358979323846 * a->rc * a->rc * a->rc;
  // (24 or more: a wave; and a block, its waves over quarters of the
  // candidates, where the other component has 2048 molecules or more)
  auto lpm = [&](int nj) { return nj * sph / vol < 24.0 ? 16 : nj >= 2048 ? 256 : 64; };
  m->lpm_a = lpm(b->n);
  m->lpm_b = lpm(a->n);
  // G6_MIX_LPM (16, 64 or 256) sets both directions' form, to test each
  if (const char *e = getenv("G6_MIX_LPM")) {
    const int f = atoi(e);
    if (f == 16 || f == 64 || f == 256) m->lpm_a = m->lpm_b = f;
  }
  auto per = [](int l) { return l == 256 ? 1 : kMixThreads / 64 * (64 / l); };
  const int per_a = per(m->lpm_a), per_b = per(m->lpm_b);
  m->nblk_a = (a->n + per_a - 1) / per_a;
  m->nblk_b = (b->n + per_b - 1) / per_b;
  if (hipMalloc(&m->part, sizeof(double) * 2 * (m->nblk_a + m->nblk_b) * kBP) != hipSuccess ||
      hipMalloc(&m->parts, sizeof(double) * 8) != hipSuccess) {
    (void)hipFree(m->part);
    delete m;
    return G6_EHIP;
  }
  G6_CHECK(hipMemset(m->parts, 0, sizeof(double) * 8));
  std::lock_guard<std::mutex> lk(g_mxmu);
  g_mix.push_back(m);
  *hm = (g6_mixture)g_mix.size();
  return 0;
}

int g6_mixture_free(g6_mixture hm) {
  Mix *m = mix_of(hm);
  if (!m) return G6_EARG;
  if (mix_live(*m)) (void)hipStreamSynchronize(m->a->stream);
  mix_drop_graphs(*m);
  (void)hipFree(m->part);
  (void)hipFree(m->parts);
  std::lock_guard<std::mutex> lk(g_mxmu);
  g_mix[hm - 1] = nullptr;
  delete m;
  return 0;
}

int g6_mixture_vreset(g6_mixture hm, double temp, uint32_t seed) {
  Mix *m = mix_of(hm);
  if (!m || !mix_live(*m) || !(temp >= 0)) return G6_EARG;
  System *s[2] = {m->a, m->b};
  return vreset_systems(s, 2, temp, seed);
}

int g6_mixture_step(g6_mixture hm, int nsteps) {
  Mix *m = mix_of(hm);
  if (!m || nsteps < 0 || !mix_live(*m)) return G6_EARG;
  System &a = *m->a, &b = *m->b;
  if (a.andersen || b.andersen) return G6_EARG;
  if (nsteps == 0) return 0;
  for (System *s : {&a, &b}) {
    const int rg = use_row_grid(*s);
    if (rg != 0) return rg;
  }
  G6_CHECK(hipStreamSynchronize(b.stream));
  if (a.graph_gen != m->gen_a || b.graph_gen != m->gen_b) {
    // a component changed what its launches embed (thermostat flag, pair
    // constants, grid)
    mix_drop_graphs(*m);
    m->gen_a = a.graph_gen;
    m->gen_b = b.graph_gen;
  }
  // the launches of the second component go to the first's stream and
  // its correctors read the first's thermostat (kernels take the buffers
  // by value at launch, so its own are back in place afterwards)
  const bool keep_na = b.nose_active;
  hipStream_t keep_st = b.stream;
  double *keep_nose = b.md.nose;
  b.nose_active = a.nose_active;
  b.stream = a.stream;
  b.md.nose = a.md.nose;
  int rc = 0;
  const int G = graph_steps();
  if (G == 0) {
    mix_steps(*m, nsteps);
  } else {
    // G-step graphs, then one of the remaining count; graphs cached by count
    auto run = [&](int k) -> int {
      int slot = -1;
      for (int q = 0; q < 4; q++)
        if (m->graph[q] && m->graph_k[q] == k) slot = q;
      if (slot < 0) {
        slot = m->graph_next;
        m->graph_next = (slot + 1) % 4;
        if (m->graph[slot]) (void)hipGraphExecDestroy(m->graph[slot]);
        m->graph[slot] = nullptr;
        hipGraph_t gr;
        if (hipStreamBeginCapture(a.stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
          return G6_EHIP;
        mix_steps(*m, k);
        if (hipStreamEndCapture(a.stream, &gr) != hipSuccess) return G6_EHIP;
        const hipError_t e = hipGraphInstantiate(&m->graph[slot], gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) {
          m->graph[slot] = nullptr;
          return G6_EHIP;
        }
        m->graph_k[slot] = k;
      }
      return hipGraphLaunch(m->graph[slot], a.stream) == hipSuccess ? 0 : G6_EHIP;
    };
    for (int k = 0; rc == 0 && k + G <= nsteps; k += G) rc = run(G);
    if (rc == 0 && nsteps % G) rc = run(nsteps % G);
  }
  b.nose_active = keep_na;
  b.stream = keep_st;
  b.md.nose = keep_nose;
  if (rc != 0) return rc;
  G6_CHECK(hipGetLastError());
  const int ra = mix_flags(a), rb = mix_flags(b);
  a.steps_done += nsteps;
  b.steps_done += nsteps;
  return ra ? ra : rb;
}

int g6_mixture_thermo(g6_mixture hm, double *thermo, double *parts) {
  Mix *m = mix_of(hm);
  if (!m || !thermo || !mix_live(*m)) return G6_EARG;
  G6_CHECK(hipStreamSynchronize(m->a->stream));
  G6_CHECK(hipMemcpy(thermo, m->a->md.thermo, sizeof(double) * G6_THERMO_LEN,
                     hipMemcpyDeviceToHost));
  if (parts) G6_CHECK(hipMemcpy(parts, m->parts, sizeof(double) * 5, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
