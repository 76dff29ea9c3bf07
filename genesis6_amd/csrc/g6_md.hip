// The MD step (the path g6_step runs), two kernels per step:
//
//   k_md_pre   64 molecules per block, one Gear role per wave, state held
//              in registers: rigid_correct with the forces of the previous
//              k_md_main, kinetic and kinetic-tensor partials, rigid_predict,
//              rigid_setmatrix, the per-molecule record read by k_md_main
//              (predicted COM in FP64, site offsets in the pair precision),
//              Address_Assign binning into fixed-capacity cell lists of this
//              step's parity, and the Verlet-skin test that decides whether
//              rows are rebuilt.
//   close      one block: reduces the block partials in a fixed order into
//              the thermo record (energies, temperature, pressure, conserved
//              H), Nose_Correct / Nose_Predict, and clears the
//              cell lists of the other parity (block 0 of k_md_main, or
//              k_md_close).  Pair partials alternate between two parities.
//   k_md_main  one wavefront per molecule i:
//              (1) if a rebuild is due, i's own neighbour row: every j != i
//                  of the cells whose box lies within rc + skin of i (a 5^3
//                  stencil of cells of edge >= (rc+skin)/2) that passes the
//                  reference's FP64 minimum-image test dd < (rc+skin)^2,
//                  sorted by j;
//              (2) the pair forces of i with every listed j whose FP64 COM
//                  distance passes dd < rc^2 - the reference's
//                  Interaction_Compress2 test, so the interacting set is the
//                  reference's - via pair_sites (g6_pair.h);
//              (3) wave reduction of i's force and lab torque (the work of
//                  rigid_collectforce), stored for the corrector;
//              (4) block partials of energy and virial.
// A step is k_md_pre(correct previous, predict) + k_md_main, whose block 0
// first does the close work of the previous step (k_md_close runs on its own
// only around the first and last step of a run); a run of steps ends with a
// correct-only k_md_pre + k_md_close, so every
// g6_step call leaves the state of the reference's loop end (after
// Nose_Correct, Mixture.F90:2282).  Cell lists and rebuild flags alternate
// between two parities so that no kernel needs a grid-wide barrier.
// Each unordered pair is evaluated in both rows, so no molecule's force
// needs an atomic, rows are sorted and every reduction runs in a fixed
// order: results are deterministic.  Energies and the virial are halved.
#include "g6_md_common.h"

// Everything written in this file is FP64 bookkeeping in the reference's
// operation order: no FMA contraction (the pragma is set by g6_md_common.h).
// The FP32 pair arithmetic lives in g6_pair.h, outside the pragma, and may
// contract.

namespace g6 {

__device__ inline float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

// force (C0 = 0) or lab torque (C0 = 3) of molecule i from the tile
// kernel's partials, summed in a fixed order (as k_part_reduce)
template <int C0>
__device__ inline void part_sum(const MdBuf &b, int i, double (&f)[3]) {
  const float *p = b.part + (size_t)i * b.nds * 6 + C0;
  for (int q = 0; q < b.nds; q++)
#pragma unroll
    for (int a = 0; a < 3; a++) f[a] += (double)p[q * 6 + a];
}

// ---- kernel A --------------------------------------------------------------
// A block updates 64 molecules (one per lane) with one role per wave:
//   wave 0  translation: com, dx[5][3]; record COM, binning, skin test
//   wave 1  angular velocity w[5][3]
//   wave 2  quaternion components a, b of every derivative; wave 3 c, d
// Every component is computed with the operation order of rigid_correct
// (RigidModule.F90:1255-1487, with the Nose drag) and rigid_predict
// (RigidModule.F90:480-559); values cross waves only through LDS, so the
// split changes no bit of the result.  The roles are wave-uniform, so no
// wave executes another role's code.
// blk is the block's index within its system (blockIdx.x for one system).
template <int BOX>
__device__ __forceinline__ void md_pre_body(const ModelTable &mt, const Geom &g, const MdBuf &b,
                                            double dt, int nose_active, int do_correct,
                                            int do_predict, int do_sites, int parity, int blk) {
  constexpr bool NPT = BOX == kBoxBaro;
  __shared__ double s_w0[3][64], s_W[12][64];
  __shared__ double s_q0[4][64], s_p[4][64], s_r[4][64], s_s[4][64], s_qn[4][64];
  const int n = g.n;
  const int role = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int i = blk * 64 + lane;
  const bool ok = i < n;
  const bool q2 = role == 2;  // wave 2: components a, b; wave 3: c, d
  double ek[kEkLen];  // Ek terms and the kinetic tensor
#pragma unroll
  for (int q = 0; q < kEkLen; q++) ek[q] = 0.0;
  double S[20];  // wave 0: com, dx; wave 1: w; waves 2-3: q (all four)
#pragma unroll
  for (int k = 0; k < 20; k++) S[k] = 0.0;
  const double zeta0 = nose_active ? b.nose[3] : 0.0;
  // translational drag: Nose, plus the Andersen dilation
  // (RigidModule.F90:1275-1305): v1 / (2*3*v0) on every axis in the
  // orthorhombic mode, v1 / (2*v0) on z alone in the orz mode
  const bool orz = NPT && b.box[16] == 2.0;
  const double acoeff = !NPT ? 0.0
                        : orz ? b.anders[3] / (2.0 * b.anders[2])
                              : b.anders[3] / (2.0 * 3.0 * b.anders[2]);
  const double nc = zeta0 * dt, exr = nc;
  const double ext[3] = {nc * 0.5 + (orz ? 0.0 : acoeff), nc * 0.5 + (orz ? 0.0 : acoeff),
                         nc * 0.5 + acoeff};
  const BoxNow bx = box_now<BOX>(g, b);
  // every global load is issued here, ahead of the first barrier
  double F[3] = {0, 0, 0}, t[9], tl[3] = {0, 0, 0}, cref[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < 9; a++) t[a] = 0.0;
  if (ok) {
    if (role == 0) {
#pragma unroll
      for (int a = 0; a < 3; a++) S[a] = b.com[a * n + i];
#pragma unroll
      for (int k = 0; k < 15; k++) S[3 + k] = b.dx[k * n + i];
      if (do_correct) {
        if (b.part) part_sum<0>(b, i, F);
        else
#pragma unroll
          for (int a = 0; a < 3; a++) F[a] = b.frc[a * n + i];
      }
      if (do_sites && b.half_skin2 > 0.0)
#pragma unroll
        for (int a = 0; a < 3; a++) cref[a] = b.com_ref[a * n + i];
    } else if (role == 1) {
#pragma unroll
      for (int k = 0; k < 15; k++) S[k] = b.w[k * n + i];
      if (do_correct) {
#pragma unroll
        for (int a = 0; a < 9; a++) t[a] = b.tmat[a * n + i];
        if (b.part) part_sum<3>(b, i, tl);
        else
#pragma unroll
          for (int a = 0; a < 3; a++) tl[a] = b.frc[(3 + a) * n + i];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 20; k++) S[k] = b.quat[k * n + i];
    }
  }
  // ---- corrector ----
  if (do_correct) {
    if (role == 1) {
#pragma unroll
      for (int a = 0; a < 3; a++) s_w0[a][lane] = S[a];  // the old w, for nw
    }
    __syncthreads();
    if (role == 0) {
      if (ok) {
        const double tf = dt * dt * 0.5 * mt.inv_mass;
#pragma unroll
        for (int a = 0; a < 3; a++) {
          double nf = F[a] * tf;
          nf = nf - S[3 + a] * ext[a];
          const double d = nf - S[6 + a];
          S[a] = S[a] + kG5_0 * d;
          S[3 + a] = S[3 + a] + kG5_1 * d;
          S[6 + a] = nf;
          S[9 + a] = S[9 + a] + kG5_3 * d;
          S[12 + a] = S[12 + a] + kG5_4 * d;
          S[15 + a] = S[15 + a] + kG5_5 * d;
        }
        ek[0] = S[3] * S[3] + S[4] * S[4] + S[5] * S[5];
        // Rigid_PressureTensor sums (RigidModule.F90:600-619), upper triangle
        ek[4] = S[3] * S[3];
        ek[5] = S[3] * S[4];
        ek[6] = S[3] * S[5];
        ek[7] = S[4] * S[4];
        ek[8] = S[4] * S[5];
        ek[9] = S[5] * S[5];
      }
    } else if (role == 1) {
      if (ok) {
        // rigid_collectforce's body-frame torque (RigidModule.F90:301-306)
        const double tq0 = t[0] * tl[0] + t[3] * tl[1] + t[6] * tl[2];
        const double tq1 = t[1] * tl[0] + t[4] * tl[1] + t[7] * tl[2];
        const double tq2 = t[2] * tl[0] + t[5] * tl[1] + t[8] * tl[2];
        const double wx = S[0], wy = S[1], wz = S[2];
        double nf[3];
        nf[0] = (tq0 + (mt.Iyy - mt.Izz) * wy * wz) * dt * mt.inv_ixx;
        nf[1] = (tq1 + (mt.Izz - mt.Ixx) * wz * wx) * dt * mt.inv_iyy;
        nf[2] = (tq2 + (mt.Ixx - mt.Iyy) * wx * wy) * dt * mt.inv_izz;
#pragma unroll
        for (int a = 0; a < 3; a++) {
          nf[a] = nf[a] - S[a] * exr;
          const double d = nf[a] - S[3 + a];
          S[a] = S[a] + kG4_0 * d;
          S[3 + a] = nf[a];
          S[6 + a] = S[6 + a] + kG4_2 * d;
          S[9 + a] = S[9 + a] + kG4_3 * d;
          S[12 + a] = S[12 + a] + kG4_4 * d;
        }
        ek[1] = S[0] * S[0];
        ek[2] = S[1] * S[1];
        ek[3] = S[2] * S[2];
      }
#pragma unroll
      for (int k = 0; k < 12; k++) s_W[k][lane] = S[k];
    } else {
      const double wx = s_w0[0][lane], wy = s_w0[1][lane], wz = s_w0[2][lane];
      const double qa = S[0], qb = S[1], qc = S[2], qd = S[3];
      if (q2) {
        const double nwa = (-wx * qb + wy * qc - wz * qd) * dt * 0.5;
        const double nwb = (wx * qa - wy * qd - wz * qc) * dt * 0.5;
        S[0] = S[0] + kG4_0 * (nwa - S[4]);
        S[1] = S[1] + kG4_0 * (nwb - S[5]);
      } else {
        const double nwc = (-wx * qd - wy * qa + wz * qb) * dt * 0.5;
        const double nwd = (wx * qc + wy * qb + wz * qa) * dt * 0.5;
        S[2] = S[2] + kG4_0 * (nwc - S[6]);
        S[3] = S[3] + kG4_0 * (nwd - S[7]);
      }
      if (q2) {
        s_q0[0][lane] = S[0];
        s_q0[1][lane] = S[1];
      } else {
        s_q0[2][lane] = S[2];
        s_q0[3][lane] = S[3];
      }
    }
    __syncthreads();
    // quaternion derivatives from the corrected w: own components, then
    // the partner's through LDS
    double w1x = 0, w1y = 0, w1z = 0, w2x = 0, w2y = 0, w2z = 0, w3x = 0, w3y = 0, w3z = 0,
           w4x = 0, w4y = 0, w4z = 0, qa = 0, qb = 0, qc = 0, qd = 0;
    if (role >= 2) {
      w1x = s_W[0][lane]; w1y = s_W[1][lane]; w1z = s_W[2][lane];
      w2x = s_W[3][lane]; w2y = s_W[4][lane]; w2z = s_W[5][lane];
      w3x = s_W[6][lane]; w3y = s_W[7][lane]; w3z = s_W[8][lane];
      w4x = s_W[9][lane]; w4y = s_W[10][lane]; w4z = s_W[11][lane];
      S[0] = s_q0[0][lane]; S[1] = s_q0[1][lane]; S[2] = s_q0[2][lane]; S[3] = s_q0[3][lane];
      qa = S[0]; qb = S[1]; qc = S[2]; qd = S[3];
      if (q2) {
        s_p[0][lane] = (-w1x * qb + w1y * qc - w1z * qd) * dt * 0.5;
        s_p[1][lane] = (w1x * qa - w1y * qd - w1z * qc) * dt * 0.5;
      } else {
        s_p[2][lane] = (-w1x * qd - w1y * qa + w1z * qb) * dt * 0.5;
        s_p[3][lane] = (w1x * qc + w1y * qb + w1z * qa) * dt * 0.5;
      }
    }
    __syncthreads();
    double pa = 0, pb = 0, pc = 0, pd = 0;
    if (role >= 2) {
      pa = s_p[0][lane]; pb = s_p[1][lane]; pc = s_p[2][lane]; pd = s_p[3][lane];
      if (q2) {
        s_r[0][lane] = (-w2x * qb + w2y * qc - w2z * qd - w1x * pb + w1y * pc - w1z * pd) * dt * 0.25;
        s_r[1][lane] = (w2x * qa - w2y * qd - w2z * qc + w1x * pa - w1y * pd - w1z * pc) * dt * 0.25;
      } else {
        s_r[2][lane] = (-w2x * qd - w2y * qa + w2z * qb - w1x * pd - w1y * pa + w1z * pb) * dt * 0.25;
        s_r[3][lane] = (w2x * qc + w2y * qb + w2z * qa + w1x * pc + w1y * pb + w1z * pa) * dt * 0.25;
      }
    }
    __syncthreads();
    double ra = 0, rb = 0, rc_ = 0, rd = 0;
    if (role >= 2) {
      ra = s_r[0][lane]; rb = s_r[1][lane]; rc_ = s_r[2][lane]; rd = s_r[3][lane];
      if (q2) {
        s_s[0][lane] = (-w3x * qb + w3y * qc - w3z * qd - w1x * rb + w1y * rc_ - w1z * rd - w2x * pb +
                        w2y * pc - w2z * pd) * dt * (1.0 / 6.0);
        s_s[1][lane] = (w3x * qa - w3y * qd - w3z * qc + w1x * ra - w1y * rd - w1z * rc_ + w2x * pa -
                        w2y * pd - w2z * pc) * dt * (1.0 / 6.0);
      } else {
        s_s[2][lane] = (-w3x * qd - w3y * qa + w3z * qb - w1x * rd - w1y * ra + w1z * rb - w2x * pd -
                        w2y * pa + w2z * pb) * dt * (1.0 / 6.0);
        s_s[3][lane] = (w3x * qc + w3y * qb + w3z * qa + w1x * rc_ + w1y * rb + w1z * ra + w2x * pc +
                        w2y * pb + w2z * pa) * dt * (1.0 / 6.0);
      }
    }
    __syncthreads();
    if (role >= 2) {
      const double sa = s_s[0][lane], sb = s_s[1][lane], sc = s_s[2][lane], sd = s_s[3][lane];
      S[4] = pa; S[5] = pb; S[6] = pc; S[7] = pd;
      S[8] = ra; S[9] = rb; S[10] = rc_; S[11] = rd;
      S[12] = sa; S[13] = sb; S[14] = sc; S[15] = sd;
      if (q2) {
        S[16] = (-w4x * qb + w4y * qc - w4z * qd - w3x * pb + w3y * pc - w3z * pd - w2x * rb +
                 w2y * rc_ - w2z * rd - w1x * sb + w1y * sc - w1z * sd) * dt * (1.0 / 8.0);
        S[17] = (w4x * qa - w4y * qd - w4z * qc + w3x * pa - w3y * pd - w3z * pc + w2x * ra -
                 w2y * rd - w2z * rc_ + w1x * sa - w1y * sd - w1z * sc) * dt * (1.0 / 8.0);
      } else {
        S[18] = (-w4x * qd - w4y * qa + w4z * qb - w3x * pd - w3y * pa + w3z * pb - w2x * rd -
                 w2y * ra + w2z * rb - w1x * sd - w1y * sa + w1z * sb) * dt * (1.0 / 8.0);
        S[19] = (w4x * qc + w4y * qb + w4z * qa + w3x * pc + w3y * pb + w3z * pa + w2x * rc_ +
                 w2y * rb + w2z * ra + w1x * sc + w1y * sb + w1z * sa) * dt * (1.0 / 8.0);
      }
    }
  }
  // ---- predictor ----
  if (do_predict) {
    if (role == 0) {
#pragma unroll
      for (int a = 0; a < 3; a++) {
        S[a] = S[a] + S[3 + a] + S[6 + a] + S[9 + a] + S[12 + a] + S[15 + a];
        S[3 + a] = S[3 + a] + 2.0 * S[6 + a] + 3.0 * S[9 + a] + 4.0 * S[12 + a] + 5.0 * S[15 + a];
        S[6 + a] = S[6 + a] + 3.0 * S[9 + a] + 6.0 * S[12 + a] + 10.0 * S[15 + a];
        S[9 + a] = S[9 + a] + 4.0 * S[12 + a] + 10.0 * S[15 + a];
        S[12 + a] = S[12 + a] + 5.0 * S[15 + a];
      }
    } else if (role == 1) {
#pragma unroll
      for (int a = 0; a < 3; a++) {
        S[a] = S[a] + S[3 + a] + S[6 + a] + S[9 + a] + S[12 + a];
        S[3 + a] = S[3 + a] + 2.0 * S[6 + a] + 3.0 * S[9 + a] + 4.0 * S[12 + a];
        S[6 + a] = S[6 + a] + 3.0 * S[9 + a] + 6.0 * S[12 + a];
        S[9 + a] = S[9 + a] + 4.0 * S[12 + a];
      }
    } else {
      if (q2) {
        predict_qc<0>(S);
        predict_qc<1>(S);
        s_qn[0][lane] = S[0];
        s_qn[1][lane] = S[1];
      } else {
        predict_qc<2>(S);
        predict_qc<3>(S);
        s_qn[2][lane] = S[2];
        s_qn[3][lane] = S[3];
      }
    }
    __syncthreads();
    if (role >= 2) {
      // all four predicted components, normalised in the order of rigid_predict
      S[0] = s_qn[0][lane]; S[1] = s_qn[1][lane]; S[2] = s_qn[2][lane]; S[3] = s_qn[3][lane];
      const double re = 1.0 / sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + S[3] * S[3]);
#pragma unroll
      for (int a = 0; a < 4; a++) S[a] = S[a] * re;
    }
  }
  if (ok && (do_predict || do_correct)) {
    if (role == 0) {
#pragma unroll
      for (int a = 0; a < 3; a++) b.com[a * n + i] = S[a];
#pragma unroll
      for (int k = 0; k < 15; k++) b.dx[k * n + i] = S[3 + k];
    } else if (role == 1) {
#pragma unroll
      for (int k = 0; k < 15; k++) b.w[k * n + i] = S[k];
    } else if (q2) {
      store_qc<0>(b, n, i, S);
    } else {
      store_qc<2>(b, n, i, S);
    }
  }
  if (do_sites && ok) {
    if (role == 0) {
      double *rec = b.rec + (size_t)i * b.rs;
      rec[0] = S[0];
      rec[1] = S[1];
      rec[2] = S[2];
      int jc[3], nd[3];
      grid_nd<BOX>(g, nd);
#pragma unroll
      for (int a = 0; a < 3; a++) {
        double d = S[a];
        const double o = round(d * bx.invL[a]) * bx.L[a];
        d = d - o;
        int j = (int)round(d * bx.cellinv[a]);
        if (j < 0) j += nd[a];
        j = ((j % nd[a]) + nd[a]) % nd[a];
        jc[a] = j;
      }
      const int c = (jc[2] * nd[1] + jc[1]) * nd[0] + jc[0];
      b.cell[i] = c;
      // slot 0 of a cell row is its count; residents follow
      int *cr = b.cellres + ((size_t)parity * g.ncell + c) * b.cell_cap;
      const int slot = atomicAdd(cr, 1) + 1;
      if (slot < b.cell_cap)
        cr[slot] = i;
      else
        raise_flag(b.flags, 3);  // cell capacity exceeded
      if (b.half_skin2 > 0.0) {
        // a pair of the rows moves by at most twice the largest displacement
        // plus |L - L_ref| (image shifts of one box per axis)
        double h2 = b.half_skin2;
        if (NPT) {
          const double e0 = bx.L[0] - b.box[9], e1 = bx.L[1] - b.box[10], e2 = bx.L[2] - b.box[11];
          const double h = sqrt(b.half_skin2) - 0.5 * sqrt(e0 * e0 + e1 * e1 + e2 * e2);
          h2 = h > 0.0 ? h * h : -1.0;
        }
        const double e0 = S[0] - cref[0], e1 = S[1] - cref[1], e2 = S[2] - cref[2];
        if (e0 * e0 + e1 * e1 + e2 * e2 > h2) b.flags[2 + parity] = 1;
      } else if (i == 0) {
        b.flags[2 + parity] = 1;  // no skin: the reference rebuilds every step
      }
      if (NPT && i == 0 &&
          (bx.L[0] < b.box[12] * b.box[15] || bx.L[1] < b.box[13] * b.box[15] ||
           bx.L[2] < b.box[14] * b.box[15]))
        raise_flag(b.flags, 5);  // cells below (rc+skin)/2
    } else if (role >= 2) {
      // rigid_setmatrix and the site offsets; the two waves split the stores
      const double qa = S[0], qb = S[1], qc = S[2], qd = S[3];
      t[0] = (qa * qa + qb * qb - (qc * qc + qd * qd));
      t[1] = -2.0 * (qa * qd + qb * qc);
      t[2] = 2.0 * (qb * qd - qa * qc);
      t[3] = 2.0 * (qa * qd - qb * qc);
      t[4] = qa * qa + qc * qc - (qb * qb + qd * qd);
      t[5] = -2.0 * (qa * qb + qc * qd);
      t[6] = 2.0 * (qa * qc + qb * qd);
      t[7] = 2.0 * (qa * qb - qc * qd);
      t[8] = qa * qa + qd * qd - (qb * qb + qc * qc);
      if (q2) {
#pragma unroll
        for (int a = 0; a < 5; a++) b.tmat[a * n + i] = t[a];
      } else {
#pragma unroll
        for (int a = 5; a < 9; a++) b.tmat[a * n + i] = t[a];
      }
      double *rec = b.rec + (size_t)i * b.rs;
      float *recf = reinterpret_cast<float *>(rec + 4);
      const int nq = mt.nsite - 1, half = (nq + 1) / 2;
      const int s0 = q2 ? 0 : half, s1 = q2 ? half : nq;
      for (int s = s0; s < s1; s++) {
        const double r0 = t[0] * mt.bx[s] + t[1] * mt.by[s] + t[2] * mt.bz[s];
        const double r1 = t[3] * mt.bx[s] + t[4] * mt.by[s] + t[5] * mt.bz[s];
        const double r2 = t[6] * mt.bx[s] + t[7] * mt.by[s] + t[8] * mt.bz[s];
        if (b.rs == kRecF) {
          recf[s * 3 + 0] = (float)r0;
          recf[s * 3 + 1] = (float)r1;
          recf[s * 3 + 2] = (float)r2;
        } else {
          rec[4 + s * 3 + 0] = r0;
          rec[4 + s * 3 + 1] = r1;
          rec[4 + s * 3 + 2] = r2;
        }
      }
    }
  }
  // kinetic partials of this wave, reduced by the step close
  // each wave reduces the partials its role produced (the rest are zero)
  if (do_correct && role == 0) {
    ek[0] = wsum63(ek[0]);
#pragma unroll
    for (int q = 4; q < kEkLen; q++) ek[q] = wsum63(ek[q]);
  } else if (do_correct && role == 1) {
#pragma unroll
    for (int q = 1; q < 4; q++) ek[q] = wsum63(ek[q]);
  }
  if (lane == 63) {
    const size_t wv = (size_t)blk * 4 + role;
#pragma unroll
    for (int q = 0; q < kEkLen; q++) b.blockek[wv * kEkLen + q] = ek[q];
  }
}

template <int BOX>
__global__ __launch_bounds__(256) void k_md_pre(ModelTable mt, Geom g, MdBuf b, double dt,
                                                int nose_active, int do_correct, int do_predict,
                                                int do_sites, int parity) {
  md_pre_body<BOX>(mt, g, b, dt, nose_active, do_correct, do_predict, do_sites, parity,
                   blockIdx.x);
}

// Ladder form: one launch over the blocks of every system of a ladder;
// map[blockIdx.x] names the system and its block (system < 0: an idle
// block that pads the XCD interleave).
__global__ __launch_bounds__(256) void k_md_pre_lad(ModelTable mt, const MdRep *__restrict__ reps,
                                                    const int2 *__restrict__ map, int do_correct,
                                                    int do_predict, int do_sites, int parity) {
  const int2 m = map[blockIdx.x];
  if (m.x < 0) return;
  const Geom g = reps[m.x].g;
  const MdBuf b = reps[m.x].b;
  md_pre_body<kBoxCube>(mt, g, b, reps[m.x].dt, reps[m.x].nose_active, do_correct, do_predict,
                     do_sites,
              parity, m.y);
}

template <bool NPT>
__global__ __launch_bounds__(256) void k_md_close(ModelTable mt, MdBuf b, int n, double dt,
                                                  int nblk_pre, const double *bpart,
                                                  int nblk_main, int ncell, int nose_active,
                                                  int mode, int parity, double box_vol) {
  __shared__ double s_red[4][kRed];
  close_body<NPT>(mt, b, n, dt, nblk_pre, bpart, nblk_main, ncell, nose_active, mode, parity,
                  box_vol, s_red);
}

// Ladder form: block r closes system r.
__global__ __launch_bounds__(256) void k_md_close_lad(ModelTable mt, const MdRep *__restrict__ reps,
                                                      int mode, int parity, int src_parity) {
  __shared__ double s_red[4][kRed];
  const MdRep &r = reps[blockIdx.x];
  const MdBuf b = r.b;
  const int n = r.g.n, nm = r.nblk_main;
  close_body<false>(mt, b, n, r.dt, (n + 63) / 64 * 4, b.blockpart + (size_t)src_parity * nm * kBP, nm,
             r.g.ncell, r.nose_active, mode, parity, box_volume(r.g.L), s_red);
}

// ---- kernel B --------------------------------------------------------------

// Cell stencil of the MD grid: per axis na offsets from lo; on an axis of
// nd >= 5 cells the 5 offsets -2..2 are distinct cells and are pruned by
// their distance from the molecule, on a shorter axis every cell is taken
// once (no pruning).
struct Stencil {
  int na[3], lo[3], nst;
  bool prune[3], any;
};
__device__ inline Stencil stencil_of(const int (&nd)[3]) {
  Stencil st;
  st.nst = 1;
  st.any = false;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    st.prune[a] = nd[a] >= 5;
    st.na[a] = st.prune[a] ? 5 : nd[a];
    st.lo[a] = st.prune[a] ? -2 : -((nd[a] - 1) / 2);
    st.nst *= st.na[a];
    st.any = st.any || st.prune[a];
  }
  return st;
}

// Distance in cell units from a point at u (cell units from the centre of
// its cell, |u| <= 1/2) to the cell at offset o, which spans [o-1/2, o+1/2];
// shrunk by 1e-9 so that rounding in the binning never drops a cell.
__device__ inline double cell_gap(int o, double u) {
  const double a = (double)o - 0.5 - u, c = u - (double)o - 0.5;
  return fmax(0.0, fmax(a, c) - 1e-9);
}

__device__ inline double frac_in_cell(const BoxNow &bx, double x, int a) {
  const double d = x - round(x * bx.invL[a]) * bx.L[a];
  const double t = d * bx.cellinv[a];
  return t - round(t);
}

// neighbour-row build of molecule i into nbr[i*maxnb ...], sorted by j so
// that the force sums do not depend on the order in which the binning
// atomics filled the cells; returns the row length (also stored).
template <int BOX>
__device__ inline int build_row(const Geom &g, const MdBuf &b, int parity, int i, int lane,
                                int (&s_beg)[128], int (&s_pre)[132], int *s_j, int *s_c) {
  constexpr bool NPT = BOX == kBoxBaro;
  const BoxNow bx = box_now<BOX>(g, b);
  const int n = g.n;
  int nd[3];
  grid_nd<BOX>(g, nd);
  const int nx = nd[0], ny = nd[1], nz = nd[2];
  const int c = b.cell[i];
  const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
  const Stencil st = stencil_of(nd);
  const double *reci = b.rec + (size_t)i * b.rs;
  const double xi = reci[0], yi = reci[1], zi = reci[2];
  const double ux = frac_in_cell(bx, xi, 0), uy = frac_in_cell(bx, yi, 1),
               uz = frac_in_cell(bx, zi, 2);
  const int *res = b.cellres + (size_t)parity * g.ncell * b.cell_cap;
  int carry = 0;
  for (int e0 = 0; e0 < st.nst; e0 += 64) {
    const int e = e0 + lane;
    int len = 0;
    if (e < st.nst) {
      const int ox = st.lo[0] + e % st.na[0], oy = st.lo[1] + (e / st.na[0]) % st.na[1],
                oz = st.lo[2] + e / (st.na[0] * st.na[1]);
      bool keep = true;
      if (BOX == kBoxCube && st.any) {  // in cell units of the cube
        const double gx = cell_gap(ox, ux), gy = cell_gap(oy, uy), gz = cell_gap(oz, uz);
        keep = gx * gx + gy * gy + gz * gz <= g.rr * (bx.cellinv[0] * bx.cellinv[0]);
      } else if (st.any) {  // the gap to the cell in A on the pruned axes
        const double gx = st.prune[0] ? cell_gap(ox, ux) / bx.cellinv[0] : 0.0;
        const double gy = st.prune[1] ? cell_gap(oy, uy) / bx.cellinv[1] : 0.0;
        const double gz = st.prune[2] ? cell_gap(oz, uz) / bx.cellinv[2] : 0.0;
        keep = gx * gx + gy * gy + gz * gz <= g.rr;
      }
      const int c2 = ((((cz + oz) % nz + nz) % nz) * ny + ((cy + oy) % ny + ny) % ny) * nx +
                     ((cx + ox) % nx + nx) % nx;
      s_beg[e] = c2 * b.cell_cap + 1;
      if (keep) len = min(res[(size_t)c2 * b.cell_cap], b.cell_cap - 1);
    }
    int inc = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int v = __shfl_up(inc, d, 64);
      if (lane >= d) inc += v;
    }
    if (e < st.nst) s_pre[e + 1] = carry + inc;
    carry += __shfl(inc, 63, 64);
  }
  if (lane == 0) s_pre[0] = 0;
  wsync();
  const int total = s_pre[st.nst];
  int base = 0;
  const int cap = b.maxnb;
  for (int t0 = 0; t0 < total; t0 += 64) {
    const int t = t0 + lane;
    bool pass = false;
    int j = 0, code = 0;
    if (t < total) {
      int lo = 0, hi = st.nst;  // s_pre[lo] <= t < s_pre[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_pre[mid] <= t) lo = mid;
        else hi = mid;
      }
      j = res[s_beg[lo] + (t - s_pre[lo])];
      if (j != i) {
        const double *recj = b.rec + (size_t)j * b.rs;
        double d0 = xi - recj[0], d1 = yi - recj[1], d2 = zi - recj[2];
        const double k0 = round(d0 * bx.invL[0]), k1 = round(d1 * bx.invL[1]),
                     k2 = round(d2 * bx.invL[2]);
        d0 = d0 - k0 * bx.L[0];
        d1 = d1 - k1 * bx.L[1];
        d2 = d2 - k2 * bx.L[2];
        const double dd = d0 * d0 + d1 * d1 + d2 * d2;
        pass = dd < g.rr;
        code = ((int)k0 + 512) | (((int)k1 + 512) << 10) | (((int)k2 + 512) << 20);
        if (pass && (fabs(k0) > 511.0 || fabs(k1) > 511.0 || fabs(k2) > 511.0)) raise_flag(b.flags, 2);
      }
    }
    const unsigned long long mask = __ballot(pass);
    const int pos = base + __popcll(mask & ((1ull << lane) - 1ull));
    if (pass) {
      if (pos < cap) {
        s_j[pos] = j;
        s_c[pos] = code;
      } else {
        raise_flag(b.flags, 4);  // neighbour-row capacity exceeded
      }
    }
    base += __popcll(mask);
  }
  wsync();
  const int m = min(base, cap);
  // rank sort by j (each j occurs once in a row)
  int *nj = b.nbr_j + (size_t)i * cap;
  int *nc = b.nbr_code + (size_t)i * cap;
  for (int t = lane; t < m; t += 64) {
    const int key = s_j[t];
    int rank = 0;
    for (int u = 0; u < m; u++) rank += s_j[u] < key;
    nj[rank] = key;
    nc[rank] = s_c[t];
  }
  wsync();  // the force loop reads entries other lanes wrote
  if (lane == 0) {
    b.nbr_cnt[i] = m;
    b.com_ref[i] = xi;
    b.com_ref[n + i] = yi;
    b.com_ref[2 * n + i] = zi;
    if (NPT && i == 0) {  // the box the rows were built in
      b.box[9] = bx.L[0];
      b.box[10] = bx.L[1];
      b.box[11] = bx.L[2];
    }
  }
  return m;
}

// blk is the block's index within its system and nblk the system's block
// count (blockIdx.x and gridDim.x for one system).
template <int M, typename R, int BOX>
__device__ __forceinline__ void md_main_body(const ModelTable &mt, const Geom &g, const MdBuf &b,
                                             const PairK<R> &pk, double *out, int parity,
                                             const CloseArgs &ca, int reg_bytes, int blk,
                                             int nblk) {
  using T = Traits<M>;
  extern __shared__ __align__(16) unsigned char s_dyn[];  // [4][reg_bytes]
  __shared__ int s_beg[4][128], s_pre[4][132];
  __shared__ double s_cnt[4][2];
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int n = g.n;
  const int i = blk * 4 + wid;
  const int cap = b.maxnb;
  const BoxNow bl = box_now<BOX>(g, b);
  unsigned char *reg = s_dyn + (size_t)wid * reg_bytes;
  double ep = 0, v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  int npair = 0, nlist = 0;  // wave-uniform counts
  if (i < n) {
    const int *nj = b.nbr_j + (size_t)i * cap;
    const int *nc = b.nbr_code + (size_t)i * cap;
    constexpr int kIdle = 512 | (512 << 10) | (512 << 20);  // zero shift
    // the first round of row entries is fetched with the row length (or
    // right after a rebuild)
    int jj[4], cc[4];
    int m;
    if (b.flags[2 + parity]) {
      int *sj = reinterpret_cast<int *>(reg);
      m = build_row<BOX>(g, b, parity, i, lane, s_beg[wid], s_pre[wid], sj, sj + cap);
      nlist = m;
    } else {
      m = b.nbr_cnt[i];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int k = 64 * u + lane;
      jj[u] = k < cap ? nj[k] : i;
      cc[u] = k < cap ? nc[k] : kIdle;
    }
    // entries at or past the row length hold stale or never-written values:
    // they are replaced before any of them forms an address
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (64 * u + lane >= m) {
        jj[u] = i;
        cc[u] = kIdle;
      }
    }
    const double *reci = b.rec + (size_t)i * b.rs;
    const double xi = reci[0], yi = reci[1], zi = reci[2];
    // pass 1: the reference's FP64 interaction test on every row entry; the
    // entries that pass are compacted in row order into LDS
    int *cj = reinterpret_cast<int *>(reg);
    R *dv = reinterpret_cast<R *>(reg + (4 * cap + 15) / 16 * 16);
    const int dstride = (int)(((size_t)cap * sizeof(R) + 15) / 16 * 16 / sizeof(R));
    wsync();  // the sort buffers of build_row share this region
    int np = 0;
    // four entries per lane per round: the index loads, then the record
    // loads, are in flight together
    for (int k0 = 0; k0 < m; k0 += 256) {
      if (k0 > 0) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int k = k0 + 64 * u + lane;
          jj[u] = k < m ? nj[k] : i;
          cc[u] = k < m ? nc[k] : kIdle;
        }
      }
      double xj[4], yj[4], zj[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const double *recj = b.rec + (size_t)jj[u] * b.rs;
        xj[u] = recj[0];
        yj[u] = recj[1];
        zj[u] = recj[2];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int code = cc[u];
        const double ox = (double)((code & 1023) - 512) * bl.L[0];
        const double oy = (double)(((code >> 10) & 1023) - 512) * bl.L[1];
        const double oz = (double)(((code >> 20) & 1023) - 512) * bl.L[2];
        const double d0 = (xi - xj[u]) - ox;
        const double d1 = (yi - yj[u]) - oy;
        const double d2 = (zi - zj[u]) - oz;
        const bool pass = k0 + 64 * u + lane < m && d0 * d0 + d1 * d1 + d2 * d2 < g.rc2;
        const unsigned long long mask = __ballot(pass);
        if (pass) {
          const int pos = np + __popcll(mask & ((1ull << lane) - 1ull));
          cj[pos] = jj[u];
          dv[pos] = R(d0);
          dv[dstride + pos] = R(d1);
          dv[2 * dstride + pos] = R(d2);
        }
        np += __popcll(mask);
      }
    }
    npair = np;
    wsync();
    // pass 2: the interacting pairs, densely, two per lane per round
    R ri[T::NS][3];
    RecSites<M, R>::load(reci, ri);
    // each lane's own pairs (two or three in a liquid) are summed in the pair
    // precision; the lane totals then go to FP64 for the wave and block sums
    R Fr[3] = {R(0), R(0), R(0)}, Tr[3] = {R(0), R(0), R(0)}, er = R(0);
    R vr[9];
#pragma unroll
    for (int q = 0; q < 9; q++) vr[q] = R(0);
    auto eval = [&](int j, R d0, R d1, R d2, const R (&rj)[T::NS][3]) {
      R Fp[3], Tp[3];
      const R e = pair_on_i<M, R>(pk, d0 * d0 + d1 * d1 + d2 * d2, d0, d1, d2, ri, rj, Fp, Tp);
      Fr[0] += Fp[0]; Fr[1] += Fp[1]; Fr[2] += Fp[2];
      Tr[0] += Tp[0]; Tr[1] += Tp[1]; Tr[2] += Tp[2];
      er += e;
      // pair virial d (x) F, fused multiply-adds
      const R dd3[3] = {d0, d1, d2};
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) vr[3 * a + c] = fma_r(dd3[a], Fp[c], vr[3 * a + c]);
    };
    for (int k0 = 0; k0 < np; k0 += 128) {
      const int ka = k0 + lane, kb = k0 + 64 + lane;
      const int ja = ka < np ? cj[ka] : i, jb = kb < np ? cj[kb] : i;
      R ra[T::NS][3], rb[T::NS][3];
      RecSites<M, R>::load(b.rec + (size_t)ja * b.rs, ra);
      RecSites<M, R>::load(b.rec + (size_t)jb * b.rs, rb);
      if (ka < np) eval(ja, dv[ka], dv[dstride + ka], dv[2 * dstride + ka], ra);
      if (kb < np) eval(jb, dv[kb], dv[dstride + kb], dv[2 * dstride + kb], rb);
    }
    double F[3], Ti[3];
    ep = (double)er;
#pragma unroll
    for (int q = 0; q < 9; q++) v[q] = (double)vr[q];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      F[a] = wsum63((double)Fr[a]);
      Ti[a] = wsum63((double)Tr[a]);
    }
    if (lane == 63) {
      if (out) {
        // molecule force (lab) and body torque, as rigid_collectforce
        double t[9];
        for (int a = 0; a < 9; a++) t[a] = b.tmat[a * n + i];
        out[i] = F[0];
        out[n + i] = F[1];
        out[2 * n + i] = F[2];
        out[3 * n + i] = t[0] * Ti[0] + t[3] * Ti[1] + t[6] * Ti[2];
        out[4 * n + i] = t[1] * Ti[0] + t[4] * Ti[1] + t[7] * Ti[2];
        out[5 * n + i] = t[2] * Ti[0] + t[5] * Ti[1] + t[8] * Ti[2];
      } else {
#pragma unroll
        for (int a = 0; a < 3; a++) {
          b.frc[a * n + i] = F[a];
          b.frc[(3 + a) * n + i] = Ti[a];
        }
      }
    }
    wsync();  // the compacted entries are dead; the region takes the partials
  }
  // block partials in a fixed order: each wave's lane partials of energy
  // and virial go to its LDS region, 16 lanes per row then sum 16 entries
  // each, and the row sums are combined
  double *acc = reinterpret_cast<double *>(reg);  // [10][64] of this wave
  acc[lane] = ep;
#pragma unroll
  for (int q = 0; q < 9; q++) acc[(1 + q) * 64 + lane] = v[q];
  if (lane == 0) {
    s_cnt[wid][0] = (double)npair;
    s_cnt[wid][1] = (double)nlist;
  }
  __syncthreads();
  if (threadIdx.x < 160) {
    // thread (q, seg) sums lanes seg, seg+16, ... of the 256: for each u the
    // 16 threads of a DPP row read 16 consecutive doubles (no bank conflict)
    // (compensated: TwoSum over the 16 terms and the row)
    const int q = threadIdx.x >> 4, seg = threadIdx.x & 15;
    double r = 0.0, rc = 0.0;
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int L = seg + 16 * u;
      two_sum_acc(r, rc,
                  reinterpret_cast<const double *>(s_dyn + (size_t)(L >> 6) * reg_bytes)[q * 64 + (L & 63)]);
    }
    rowsum16_dd(r, rc);
    if (seg == 0) b.blockpart[((size_t)parity * nblk + blk) * kBP + q] = 0.5 * (r + rc);
  } else if (threadIdx.x < 162) {
    const int q = threadIdx.x - 160;
    b.blockpart[((size_t)parity * nblk + blk) * kBP + 10 + q] =
        0.5 * (s_cnt[0][q] + s_cnt[1][q] + s_cnt[2][q] + s_cnt[3][q]);
  }
}

// Waves per SIMD the row kernel is compiled for (register budget)
template <int M, typename R>
constexpr int main_waves_per_eu() {
  // FP32 TIP4P and SPC/E rows fit in 96 VGPRs: five waves per SIMD, so that a
  // 4096-molecule grid plus the step-close block (1025 blocks) is resident
  // at once on 256 CUs (four waves per SIMD leave the last block waiting for
  // a slot, which set the kernel's end)
  if (sizeof(R) == 4 && (M == G6_TIP4P || M == G6_SPCE || M == G6_OPLSMEOH)) return 5;
  return (M != G6_ST2 && M != G6_NVDE && sizeof(R) == 4) ? 4 : 1;
}

template <int M, typename R, int BOX>
__global__ __launch_bounds__(256, (main_waves_per_eu<M, R>())) void k_md_main(ModelTable mt, Geom g, MdBuf b, PairK<R> pk,
                                                 double *out, int parity, CloseArgs ca,
                                                 int reg_bytes) {
  if (ca.mode) {
    // block 0 closes the previous step while the other blocks run the rows
    if (blockIdx.x == 0) {
      __shared__ double s_close[4][kRed];
      close_body<false>(mt, b, g.n, ca.dt, ca.nblk_pre, ca.bpart, ca.nblk_main, g.ncell,
                        ca.nose_active, ca.mode, parity, box_volume(g.L), s_close);
      return;
    }
    md_main_body<M, R, BOX>(mt, g, b, pk, out, parity, ca, reg_bytes, blockIdx.x - 1,
                            gridDim.x - 1);
  } else {
    md_main_body<M, R, BOX>(mt, g, b, pk, out, parity, ca, reg_bytes, blockIdx.x, gridDim.x);
  }
}

// Ladder form (map as k_md_pre_lad); the block of index -1 of each system
// closes that system's previous step when close_mode != 0.
template <int M, typename R>
__global__ __launch_bounds__(256, (main_waves_per_eu<M, R>())) void k_md_main_lad(ModelTable mt, const MdRep *__restrict__ reps,
                                                     const int2 *__restrict__ map, PairK<R> pk,
                                                     int parity, int close_mode, int reg_bytes) {
  const int2 m = map[blockIdx.x];
  if (m.x < 0) return;
  const Geom g = reps[m.x].g;
  const MdBuf b = reps[m.x].b;
  const int nm = reps[m.x].nblk_main;
  CloseArgs ca;
  ca.mode = close_mode;
  ca.nose_active = reps[m.x].nose_active;
  ca.nblk_pre = (g.n + 63) / 64 * 4;
  ca.nblk_main = nm;
  ca.bpart = b.blockpart + (size_t)(1 - parity) * nm * kBP;
  ca.dt = reps[m.x].dt;
  if (m.y < 0) {
    if (close_mode) {
      __shared__ double s_close[4][kRed];
      close_body<false>(mt, b, g.n, ca.dt, ca.nblk_pre, ca.bpart, ca.nblk_main, g.ncell,
                        ca.nose_active, ca.mode, parity, box_volume(g.L), s_close);
    }
    return;
  }
  md_main_body<M, R, kBoxCube>(mt, g, b, pk, nullptr, parity, ca, reg_bytes, m.y, nm);
}

__global__ void k_flag_set(int *flags, int k, int v) {
  if (threadIdx.x == 0) flags[k] = v;
}

void md_launch_flag(System &s, int k, int v) { k_flag_set<<<1, 64, 0, s.stream>>>(s.md.flags, k, v); }

void md_launch_pre(System &s, bool correct, bool predict, bool sites, int parity) {
  const int nb = pre_waves(s.n) / 4, na = s.nose_active ? 1 : 0;
  if (s.andersen)
    k_md_pre<kBoxBaro><<<nb, 256, 0, s.stream>>>(s.mt, s.geo_md, s.md, s.dt, na, correct ? 1 : 0,
                                                 predict ? 1 : 0, sites ? 1 : 0, parity);
  else if (is_cubic(s))
    k_md_pre<kBoxCube><<<nb, 256, 0, s.stream>>>(s.mt, s.geo_md, s.md, s.dt, na, correct ? 1 : 0,
                                                 predict ? 1 : 0, sites ? 1 : 0, parity);
  else
    k_md_pre<kBoxOrtho><<<nb, 256, 0, s.stream>>>(s.mt, s.geo_md, s.md, s.dt, na,
                                                  correct ? 1 : 0, predict ? 1 : 0, sites ? 1 : 0,
                                                  parity);
}

static CloseArgs close_args(System &s, int mode, int src_parity) {
  CloseArgs ca;
  ca.mode = mode;
  ca.nose_active = s.nose_active ? 1 : 0;
  ca.nblk_pre = pre_waves(s.n);
  ca.nblk_main = s.nblk_md;
  ca.bpart = s.md.blockpart + (size_t)src_parity * s.nblk_md * kBP;
  ca.dt = s.dt;
  return ca;
}

void md_launch_close(System &s, int mode, int parity, int src_parity) {
  const CloseArgs ca = close_args(s, mode, src_parity);
  if (s.andersen)
    k_md_close<true><<<1, 256, 0, s.stream>>>(s.mt, s.md, s.n, s.dt, ca.nblk_pre, ca.bpart,
                                              ca.nblk_main, s.geo_md.ncell, ca.nose_active, mode,
                                              parity, box_volume(s.Lb));
  else
    k_md_close<false><<<1, 256, 0, s.stream>>>(s.mt, s.md, s.n, s.dt, ca.nblk_pre, ca.bpart,
                                               ca.nblk_main, s.geo_md.ncell, ca.nose_active,
                                               mode, parity, box_volume(s.Lb));
}

template <typename R>
static const PairK<R> &pair_k(const System &s);
template <>
const PairK<float> &pair_k<float>(const System &s) { return s.pkf; }
template <>
const PairK<double> &pair_k<double>(const System &s) { return s.pkd; }

template <int M, typename R>
static void md_main_t(System &s, double *out, int parity, const CloseArgs &ca) {
  const int rb = main_region_bytes(s.md.maxnb, (int)sizeof(R));
  const unsigned grid = s.nblk_md + (ca.mode ? 1 : 0);
  const size_t lds = 4 * (size_t)rb;
  if (s.andersen)
    k_md_main<M, R, kBoxBaro><<<grid, 256, lds, s.stream>>>(s.mt, s.geo_md, s.md, pair_k<R>(s),
                                                            out, parity, ca, rb);
  else if (is_cubic(s))
    k_md_main<M, R, kBoxCube><<<grid, 256, lds, s.stream>>>(s.mt, s.geo_md, s.md, pair_k<R>(s),
                                                            out, parity, ca, rb);
  else
    k_md_main<M, R, kBoxOrtho><<<grid, 256, lds, s.stream>>>(s.mt, s.geo_md, s.md, pair_k<R>(s),
                                                             out, parity, ca, rb);
}

void md_launch_main(System &s, double *out, int parity, int close_mode) {
  const CloseArgs ca = close_args(s, close_mode, 1 - parity);
  if (s.tiles) {
    md_launch_tile(s, parity, ca);
    if (out) md_launch_part_reduce(s, out);
    return;
  }
  if (s.prec == G6_FP32) {
    if (s.kmodel == G6_TIP4P) md_main_t<G6_TIP4P, float>(s, out, parity, ca);
    else if (s.kmodel == G6_TIP5P) md_main_t<G6_TIP5P, float>(s, out, parity, ca);
    else if (s.kmodel == G6_SPCE) md_main_t<G6_SPCE, float>(s, out, parity, ca);
    else if (s.kmodel == G6_NVDE) md_main_t<G6_NVDE, float>(s, out, parity, ca);
    else if (s.kmodel == G6_OPLSMEOH) md_main_t<G6_OPLSMEOH, float>(s, out, parity, ca);
    else md_main_t<G6_ST2, float>(s, out, parity, ca);
  } else {
    if (s.kmodel == G6_TIP4P) md_main_t<G6_TIP4P, double>(s, out, parity, ca);
    else if (s.kmodel == G6_TIP5P) md_main_t<G6_TIP5P, double>(s, out, parity, ca);
    else if (s.kmodel == G6_SPCE) md_main_t<G6_SPCE, double>(s, out, parity, ca);
    else if (s.kmodel == G6_NVDE) md_main_t<G6_NVDE, double>(s, out, parity, ca);
    else if (s.kmodel == G6_OPLSMEOH) md_main_t<G6_OPLSMEOH, double>(s, out, parity, ca);
    else md_main_t<G6_ST2, double>(s, out, parity, ca);
  }
}

void md_launch_pre_lad(Ladder &l, bool correct, bool predict, bool sites, int parity) {
  k_md_pre_lad<<<l.grid_pre, 256, 0, l.stream>>>(l.mt, l.reps, l.map_pre, correct ? 1 : 0,
                                                 predict ? 1 : 0, sites ? 1 : 0, parity);
}

void md_launch_close_lad(Ladder &l, int mode, int parity, int src_parity) {
  k_md_close_lad<<<(int)l.sys.size(), 256, 0, l.stream>>>(l.mt, l.reps, mode, parity, src_parity);
}

template <int M, typename R>
static void md_main_lad_t(Ladder &l, const PairK<R> &pk, int parity, int close_mode) {
  k_md_main_lad<M, R><<<l.grid_main, 256, 4 * (size_t)l.reg_bytes, l.stream>>>(
      l.mt, l.reps, l.map_main, pk, parity, close_mode, l.reg_bytes);
}

void md_launch_main_lad(Ladder &l, int parity, int close_mode) {
  if (l.prec == G6_FP32) {
    if (l.model == G6_TIP4P) md_main_lad_t<G6_TIP4P, float>(l, l.pkf, parity, close_mode);
    else if (l.model == G6_TIP5P) md_main_lad_t<G6_TIP5P, float>(l, l.pkf, parity, close_mode);
    else if (l.model == G6_SPCE) md_main_lad_t<G6_SPCE, float>(l, l.pkf, parity, close_mode);
    else if (l.model == G6_NVDE) md_main_lad_t<G6_NVDE, float>(l, l.pkf, parity, close_mode);
    else if (l.model == G6_OPLSMEOH)
      md_main_lad_t<G6_OPLSMEOH, float>(l, l.pkf, parity, close_mode);
    else md_main_lad_t<G6_ST2, float>(l, l.pkf, parity, close_mode);
  } else {
    if (l.model == G6_TIP4P) md_main_lad_t<G6_TIP4P, double>(l, l.pkd, parity, close_mode);
    else if (l.model == G6_TIP5P) md_main_lad_t<G6_TIP5P, double>(l, l.pkd, parity, close_mode);
    else if (l.model == G6_SPCE) md_main_lad_t<G6_SPCE, double>(l, l.pkd, parity, close_mode);
    else if (l.model == G6_NVDE) md_main_lad_t<G6_NVDE, double>(l, l.pkd, parity, close_mode);
    else if (l.model == G6_OPLSMEOH)
      md_main_lad_t<G6_OPLSMEOH, double>(l, l.pkd, parity, close_mode);
    else md_main_lad_t<G6_ST2, double>(l, l.pkd, parity, close_mode);
  }
}

}  // namespace g6
