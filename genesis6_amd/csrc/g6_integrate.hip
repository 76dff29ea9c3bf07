// Per-molecule kernels of the MD step: predictor, site placement and cell
// binning (k_pre), corrector with the Nose drag and kinetic-energy partials
// (k_correct), and the one-block reduction that closes the step
// (k_thermo: temperature, Nose_Correct, conserved quantity).
//
//   rigid_predict         RigidModule.F90:480-559
//   Nose_Predict/Correct  NoseModule.F90:43-87
//   rigid_setmatrix + rigid_setsiteposition0  RigidModule.F90:62-128
//   Address_Assign binning GridModule.F90:503-536
//   rigid_correct         RigidModule.F90:1255-1487 (Nose terms only)
//   Rigid_Ek              RigidModule.F90:576-597
//   temperature, H        Mixture.F90:2251, 2431-2468
// All FP64, evaluated in the reference's operation order (no contraction).
#include "g6_internal.h"

#pragma clang fp contract(off)

namespace g6 {

constexpr double kGear5_0 = 3.0 / 16.0, kGear5_1 = 251.0 / 360.0,
                 kGear5_3 = 11.0 / 18.0, kGear5_4 = 1.0 / 6.0,
                 kGear5_5 = 1.0 / 60.0;
constexpr double kGear4_0 = 251.0 / 720.0, kGear4_2 = 11.0 / 12.0,
                 kGear4_3 = 1.0 / 3.0, kGear4_4 = 1.0 / 24.0;
constexpr double kI2J = 10.0;
#define G6_RGAS (6.0225e23 * 1.380066e-23)

__device__ inline void set_matrix(const double *quat, int n, int i, double *t) {
  const double qa = quat[0 * n + i], qb = quat[1 * n + i], qc = quat[2 * n + i],
               qd = quat[3 * n + i];
  t[0] = (qa * qa + qb * qb - (qc * qc + qd * qd));
  t[1] = -2.0 * (qa * qd + qb * qc);
  t[2] = 2.0 * (qb * qd - qa * qc);
  t[3] = 2.0 * (qa * qd - qb * qc);
  t[4] = qa * qa + qc * qc - (qb * qb + qd * qd);
  t[5] = -2.0 * (qa * qb + qc * qd);
  t[6] = 2.0 * (qa * qc + qb * qd);
  t[7] = 2.0 * (qa * qb - qc * qd);
  t[8] = qa * qa + qd * qd - (qb * qb + qc * qc);
}

__device__ inline void nose_predict(double *z) {
  z[3] = z[3] + z[4] + z[5] + z[6] + z[7];
  z[4] = z[4] + 2.0 * z[5] + 3.0 * z[6] + 4.0 * z[7];
  z[5] = z[5] + 3.0 * z[6] + 6.0 * z[7];
  z[6] = z[6] + 4.0 * z[7];
}

__global__ __launch_bounds__(256) void k_pre(ModelTable mt, Geom g, double *com,
                                             double *dx, double *w, double *quat,
                                             double *tmat, double *intra,
                                             float *intraf, int *cell, int *count,
                                             const double *com_ref, int *flags,
                                             double *nose, int nose_active,
                                             int predict, double half_skin2) {
  const int n = g.n;
  if (predict && nose_active && blockIdx.x == 0 && threadIdx.x == 0) nose_predict(nose);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (predict) {
#define DX(s, a) dx[((s) * 3 + (a)) * n + i]
#define W(s, a) w[((s) * 3 + (a)) * n + i]
#define Q(s, a) quat[((s) * 4 + (a)) * n + i]
    for (int a = 0; a < 3; a++) {
      com[a * n + i] = com[a * n + i] + DX(0, a) + DX(1, a) + DX(2, a) + DX(3, a) + DX(4, a);
      DX(0, a) = DX(0, a) + 2.0 * DX(1, a) + 3.0 * DX(2, a) + 4.0 * DX(3, a) + 5.0 * DX(4, a);
      DX(1, a) = DX(1, a) + 3.0 * DX(2, a) + 6.0 * DX(3, a) + 10.0 * DX(4, a);
      DX(2, a) = DX(2, a) + 4.0 * DX(3, a) + 10.0 * DX(4, a);
      DX(3, a) = DX(3, a) + 5.0 * DX(4, a);
    }
    for (int a = 0; a < 3; a++) {
      W(0, a) = W(0, a) + W(1, a) + W(2, a) + W(3, a) + W(4, a);
      W(1, a) = W(1, a) + 2.0 * W(2, a) + 3.0 * W(3, a) + 4.0 * W(4, a);
      W(2, a) = W(2, a) + 3.0 * W(3, a) + 6.0 * W(4, a);
      W(3, a) = W(3, a) + 4.0 * W(4, a);
    }
    for (int a = 0; a < 4; a++) {
      Q(0, a) = Q(0, a) + Q(1, a) + Q(2, a) + Q(3, a) + Q(4, a);
      Q(1, a) = Q(1, a) + 2.0 * Q(2, a) + 3.0 * Q(3, a) + 4.0 * Q(4, a);
    }
    const double re = 1.0 / sqrt(Q(0, 0) * Q(0, 0) + Q(0, 1) * Q(0, 1) +
                                 Q(0, 2) * Q(0, 2) + Q(0, 3) * Q(0, 3));
    for (int a = 0; a < 4; a++) Q(0, a) = Q(0, a) * re;
#undef DX
#undef W
#undef Q
  }
  double t[9];
  set_matrix(quat, n, i, t);
  for (int a = 0; a < 9; a++) tmat[a * n + i] = t[a];
  for (int s = 0; s < mt.nsite; s++) {
    const double r0 = t[0] * mt.bx[s] + t[1] * mt.by[s] + t[2] * mt.bz[s];
    const double r1 = t[3] * mt.bx[s] + t[4] * mt.by[s] + t[5] * mt.bz[s];
    const double r2 = t[6] * mt.bx[s] + t[7] * mt.by[s] + t[8] * mt.bz[s];
    intra[(s * 3 + 0) * n + i] = r0;
    intra[(s * 3 + 1) * n + i] = r1;
    intra[(s * 3 + 2) * n + i] = r2;
    intraf[(s * 3 + 0) * n + i] = (float)r0;
    intraf[(s * 3 + 1) * n + i] = (float)r1;
    intraf[(s * 3 + 2) * n + i] = (float)r2;
  }
  // Address_Assign: wrap, round to the centred cell, mAddr
  int jc[3];
  for (int a = 0; a < 3; a++) {
    double d = com[a * n + i];
    const double o = round(d * g.invL) * g.L;
    d = d - o;
    int j = (int)round(d * g.cellinv);
    if (j < 0) j += g.ndiv;
    j = ((j % g.ndiv) + g.ndiv) % g.ndiv;  // only acts for ndiv = 1
    jc[a] = j;
  }
  const int c = (jc[2] * g.ndiv + jc[1]) * g.ndiv + jc[0];
  cell[i] = c;
  atomicAdd(&count[c], 1);
  if (half_skin2 > 0.0) {
    const double e0 = com[i] - com_ref[i], e1 = com[n + i] - com_ref[n + i],
                 e2 = com[2 * n + i] - com_ref[2 * n + i];
    if (e0 * e0 + e1 * e1 + e2 * e2 > half_skin2) flags[1] = 1;
  }
}

void launch_predict_sites_bin(System &s, bool predict) {
  const double hs = s.skin > 0 ? 0.25 * s.skin * s.skin : 0.0;
  k_pre<<<s.nblk_mol, 256, 0, s.stream>>>(s.mt, s.geo, s.com, s.dx, s.w, s.quat, s.tmat,
                                          s.intra, s.intraf, s.cell, s.cell_count,
                                          s.com_ref, s.flags, s.nose, s.nose_active ? 1 : 0,
                                          predict ? 1 : 0, hs);
}

__global__ __launch_bounds__(256) void k_correct(ModelTable mt, Geom g, double dt,
                                                 double *com, double *dx, double *w,
                                                 double *quat, const double *tmat,
                                                 double *facc, const double *nose,
                                                 int nose_active, double *blockpart,
                                                 int *count) {
  const int n = g.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double ek[4] = {0, 0, 0, 0};
  for (int c = i; c <= g.ncell; c += gridDim.x * blockDim.x) count[c] = 0;
  if (i < n) {
    double F[3], tl[3];
    for (int a = 0; a < 3; a++) {
      F[a] = facc[a * n + i] + facc[(6 + a) * n + i];
      tl[a] = facc[(3 + a) * n + i] + facc[(9 + a) * n + i];
      facc[a * n + i] = 0.0;
      facc[(3 + a) * n + i] = 0.0;
    }
    double t[9];
    for (int a = 0; a < 9; a++) t[a] = tmat[a * n + i];
    const double tq[3] = {t[0] * tl[0] + t[3] * tl[1] + t[6] * tl[2],
                          t[1] * tl[0] + t[4] * tl[1] + t[7] * tl[2],
                          t[2] * tl[0] + t[5] * tl[1] + t[8] * tl[2]};
    const double zeta0 = nose_active ? nose[3] : 0.0;
    const double nc = zeta0 * dt, ext = nc * 0.5, exr = nc;
    const double tf = dt * dt * 0.5 * (1.0 / mt.mass);
#define DX(s, a) dx[((s) * 3 + (a)) * n + i]
#define W(s, a) w[((s) * 3 + (a)) * n + i]
#define Q(s, a) quat[((s) * 4 + (a)) * n + i]
    for (int a = 0; a < 3; a++) {
      double nf = F[a] * tf;
      nf = nf - DX(0, a) * ext;
      const double d = nf - DX(1, a);
      com[a * n + i] = com[a * n + i] + kGear5_0 * d;
      DX(0, a) = DX(0, a) + kGear5_1 * d;
      DX(1, a) = nf;
      DX(2, a) = DX(2, a) + kGear5_3 * d;
      DX(3, a) = DX(3, a) + kGear5_4 * d;
      DX(4, a) = DX(4, a) + kGear5_5 * d;
    }
    double wx = W(0, 0), wy = W(0, 1), wz = W(0, 2);
    double qa = Q(0, 0), qb = Q(0, 1), qc = Q(0, 2), qd = Q(0, 3);
    const double nwa = (-wx * qb + wy * qc - wz * qd) * dt * 0.5;
    const double nwb = (wx * qa - wy * qd - wz * qc) * dt * 0.5;
    const double nwc = (-wx * qd - wy * qa + wz * qb) * dt * 0.5;
    const double nwd = (wx * qc + wy * qb + wz * qa) * dt * 0.5;
    const double dwa = nwa - Q(1, 0), dwb = nwb - Q(1, 1), dwc = nwc - Q(1, 2),
                 dwd = nwd - Q(1, 3);
    Q(0, 0) = Q(0, 0) + kGear4_0 * dwa;
    Q(0, 1) = Q(0, 1) + kGear4_0 * dwb;
    Q(0, 2) = Q(0, 2) + kGear4_0 * dwc;
    Q(0, 3) = Q(0, 3) + kGear4_0 * dwd;
    double nf[3];
    nf[0] = (tq[0] + (mt.Iyy - mt.Izz) * wy * wz) * dt * (1.0 / mt.Ixx);
    nf[1] = (tq[1] + (mt.Izz - mt.Ixx) * wz * wx) * dt * (1.0 / mt.Iyy);
    nf[2] = (tq[2] + (mt.Ixx - mt.Iyy) * wx * wy) * dt * (1.0 / mt.Izz);
    for (int a = 0; a < 3; a++) {
      nf[a] = nf[a] - W(0, a) * exr;
      const double d = nf[a] - W(1, a);
      W(0, a) = W(0, a) + kGear4_0 * d;
      W(1, a) = nf[a];
      W(2, a) = W(2, a) + kGear4_2 * d;
      W(3, a) = W(3, a) + kGear4_3 * d;
      W(4, a) = W(4, a) + kGear4_4 * d;
    }
    const double w1x = W(0, 0), w1y = W(0, 1), w1z = W(0, 2);
    const double w2x = W(1, 0), w2y = W(1, 1), w2z = W(1, 2);
    const double w3x = W(2, 0), w3y = W(2, 1), w3z = W(2, 2);
    const double w4x = W(3, 0), w4y = W(3, 1), w4z = W(3, 2);
    qa = Q(0, 0); qb = Q(0, 1); qc = Q(0, 2); qd = Q(0, 3);
    Q(1, 2) = (-w1x * qd - w1y * qa + w1z * qb) * dt * 0.5;
    Q(1, 1) = (w1x * qa - w1y * qd - w1z * qc) * dt * 0.5;
    Q(1, 3) = (w1x * qc + w1y * qb + w1z * qa) * dt * 0.5;
    Q(1, 0) = (-w1x * qb + w1y * qc - w1z * qd) * dt * 0.5;
    const double pa = Q(1, 0), pb = Q(1, 1), pc = Q(1, 2), pd = Q(1, 3);
    Q(2, 2) = (-w2x * qd - w2y * qa + w2z * qb - w1x * pd - w1y * pa + w1z * pb) * dt * 0.25;
    Q(2, 1) = (w2x * qa - w2y * qd - w2z * qc + w1x * pa - w1y * pd - w1z * pc) * dt * 0.25;
    Q(2, 3) = (w2x * qc + w2y * qb + w2z * qa + w1x * pc + w1y * pb + w1z * pa) * dt * 0.25;
    Q(2, 0) = (-w2x * qb + w2y * qc - w2z * qd - w1x * pb + w1y * pc - w1z * pd) * dt * 0.25;
    const double ra = Q(2, 0), rb = Q(2, 1), rc = Q(2, 2), rd = Q(2, 3);
    Q(3, 2) = (-w3x * qd - w3y * qa + w3z * qb - w1x * rd - w1y * ra + w1z * rb - w2x * pd -
               w2y * pa + w2z * pb) * dt * (1.0 / 6.0);
    Q(3, 1) = (w3x * qa - w3y * qd - w3z * qc + w1x * ra - w1y * rd - w1z * rc + w2x * pa -
               w2y * pd - w2z * pc) * dt * (1.0 / 6.0);
    Q(3, 3) = (w3x * qc + w3y * qb + w3z * qa + w1x * rc + w1y * rb + w1z * ra + w2x * pc +
               w2y * pb + w2z * pa) * dt * (1.0 / 6.0);
    Q(3, 0) = (-w3x * qb + w3y * qc - w3z * qd - w1x * rb + w1y * rc - w1z * rd - w2x * pb +
               w2y * pc - w2z * pd) * dt * (1.0 / 6.0);
    const double sa = Q(3, 0), sb = Q(3, 1), sc = Q(3, 2), sd = Q(3, 3);
    Q(4, 2) = (-w4x * qd - w4y * qa + w4z * qb - w3x * pd - w3y * pa + w3z * pb - w2x * rd -
               w2y * ra + w2z * rb - w1x * sd - w1y * sa + w1z * sb) * dt * (1.0 / 8.0);
    Q(4, 1) = (w4x * qa - w4y * qd - w4z * qc + w3x * pa - w3y * pd - w3z * pc + w2x * ra -
               w2y * rd - w2z * rc + w1x * sa - w1y * sd - w1z * sc) * dt * (1.0 / 8.0);
    Q(4, 3) = (w4x * qc + w4y * qb + w4z * qa + w3x * pc + w3y * pb + w3z * pa + w2x * rc +
               w2y * rb + w2z * ra + w1x * sc + w1y * sb + w1z * sa) * dt * (1.0 / 8.0);
    Q(4, 0) = (-w4x * qb + w4y * qc - w4z * qd - w3x * pb + w3y * pc - w3z * pd - w2x * rb +
               w2y * rc - w2z * rd - w1x * sb + w1y * sc - w1z * sd) * dt * (1.0 / 8.0);
    ek[0] = DX(0, 0) * DX(0, 0) + DX(0, 1) * DX(0, 1) + DX(0, 2) * DX(0, 2);
    ek[1] = W(0, 0) * W(0, 0);
    ek[2] = W(0, 1) * W(0, 1);
    ek[3] = W(0, 2) * W(0, 2);
#undef DX
#undef W
#undef Q
  }
  // block partial sums of the kinetic terms (fixed order: deterministic)
  __shared__ double red[4][256];
  for (int q = 0; q < 4; q++) red[q][threadIdx.x] = ek[q];
  __syncthreads();
  for (int d = 128; d >= 1; d >>= 1) {
    if ((int)threadIdx.x < d)
      for (int q = 0; q < 4; q++) red[q][threadIdx.x] += red[q][threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x < 4) blockpart[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// molecule force and body torque of the current accumulators (no motion)
__global__ void k_collect_acc(Geom g, const double *tmat, double *facc, double *out) {
  const int n = g.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double tl[3];
  for (int a = 0; a < 3; a++) {
    out[a * n + i] = facc[a * n + i] + facc[(6 + a) * n + i];
    tl[a] = facc[(3 + a) * n + i] + facc[(9 + a) * n + i];
    facc[a * n + i] = 0.0;
    facc[(3 + a) * n + i] = 0.0;
  }
  double t[9];
  for (int a = 0; a < 9; a++) t[a] = tmat[a * n + i];
  out[3 * n + i] = t[0] * tl[0] + t[3] * tl[1] + t[6] * tl[2];
  out[4 * n + i] = t[1] * tl[0] + t[4] * tl[1] + t[7] * tl[2];
  out[5 * n + i] = t[2] * tl[0] + t[5] * tl[1] + t[8] * tl[2];
}

__global__ void k_zero_ek(double *blockpart, int nblk, int *count, int ncell) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = i; k < 4 * nblk; k += gridDim.x * blockDim.x) blockpart[k] = 0.0;
  for (int c = i; c <= ncell; c += gridDim.x * blockDim.x) count[c] = 0;
}

void launch_collect_acc(System &s, double *out) {
  k_collect_acc<<<s.nblk_mol, 256, 0, s.stream>>>(s.geo, s.tmat, s.facc, out);
  k_zero_ek<<<s.nblk_mol, 256, 0, s.stream>>>(s.blockpart, s.nblk_mol, s.cell_count,
                                              s.geo.ncell);
}

void launch_correct(System &s) {
  k_correct<<<s.nblk_mol, 256, 0, s.stream>>>(s.mt, s.geo, s.dt, s.com, s.dx, s.w, s.quat,
                                              s.tmat, s.facc, s.nose, s.nose_active ? 1 : 0,
                                              s.blockpart, s.cell_count);
}

// One block: sums of the row partials (energy, virial) and the kinetic block
// partials in a fixed order; temperature; Nose_Correct; H; the thermo record.
__global__ __launch_bounds__(256) void k_thermo(ModelTable mt, int n, int nblk, double dt,
                                                const double *rowpart,
                                                const double *blockpart, double *nose,
                                                int nose_active, double *thermo,
                                                int *flags, int reset_gate,
                                                const int *off) {
  constexpr int NQ = 15;
  __shared__ double red[NQ][256];
  double acc[NQ];
  for (int q = 0; q < NQ; q++) acc[q] = 0.0;
  for (int r = threadIdx.x; r < n; r += 256)
    for (int q = 0; q < 11; q++) acc[q] += rowpart[q * n + r];
  for (int b = threadIdx.x; b < nblk; b += 256)
    for (int q = 0; q < 4; q++) acc[11 + q] += blockpart[b * 4 + q];
  for (int q = 0; q < NQ; q++) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int d = 128; d >= 1; d >>= 1) {
    if ((int)threadIdx.x < d)
      for (int q = 0; q < NQ; q++) red[q][threadIdx.x] += red[q][threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double ep = red[0][0];
    const double ekt = red[11][0] * mt.mass * 0.5 * (1.0 / (dt * dt));
    const double ekr = (red[12][0] * mt.Ixx + red[13][0] * mt.Iyy + red[14][0] * mt.Izz) * 0.5;
    const double dof = 6.0 * n;
    const double temp = kI2J * 2.0 * (ekt + ekr) / (G6_RGAS * dof);
    double h = (ep + ekt + ekr) * kI2J;
    if (nose_active) {
      const double zetaf = (temp - nose[0]) / nose[1] - nose[2] * nose[3];
      const double dz2 = zetaf - nose[4];
      nose[3] = nose[3] + kGear4_0 * dz2;
      nose[4] = zetaf;
      nose[5] = nose[5] + kGear4_2 * dz2;
      nose[6] = nose[6] + kGear4_3 * dz2;
      nose[7] = nose[7] + kGear4_4 * dz2;
      nose[8] = nose[8] + nose[3];
      const double nose_ek = (nose[3] * nose[3] * nose[1] * dt * dof * 0.5) * G6_RGAS;
      const double nose_ep = (dof * nose[0] * (nose[8] - 0.5 * nose[3]) * dt) * G6_RGAS;
      h = h + nose_ep + nose_ek;
    }
    thermo[0] = ep;
    thermo[1] = ekt;
    thermo[2] = ekr;
    thermo[3] = temp;
    for (int q = 0; q < 9; q++) thermo[4 + q] = red[1 + q][0];
    thermo[13] = h;
    thermo[14] = red[10][0];
    thermo[15] = (double)off[2 * n];
    if (reset_gate) flags[1] = 0;
  }
}

void launch_thermo(System &s) {
  k_thermo<<<1, 256, 0, s.stream>>>(s.mt, s.n, s.nblk_mol, s.dt, s.rowpart, s.blockpart,
                                    s.nose, s.nose_active ? 1 : 0, s.thermo, s.flags,
                                    s.skin > 0 ? 1 : 0, s.off);
}

// ---- module seam helpers on host-given arrays ----------------------------
__global__ void k_sites(ModelTable mt, int n, const double *com, const double *quat1,
                        double *x, double *y, double *z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double t[9];
  set_matrix(quat1, n, i, t);
  const int ns = mt.nsite;
  for (int s = 0; s < ns; s++) {
    const double r0 = t[0] * mt.bx[s] + t[1] * mt.by[s] + t[2] * mt.bz[s];
    const double r1 = t[3] * mt.bx[s] + t[4] * mt.by[s] + t[5] * mt.bz[s];
    const double r2 = t[6] * mt.bx[s] + t[7] * mt.by[s] + t[8] * mt.bz[s];
    x[i * ns + s] = com[i] + r0;
    y[i * ns + s] = com[n + i] + r1;
    z[i * ns + s] = com[2 * n + i] + r2;
  }
}

__global__ void k_collect(ModelTable mt, int n, const double *quat1, const double *fx,
                          const double *fy, const double *fz, double *force,
                          double *torque) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double t[9];
  set_matrix(quat1, n, i, t);
  const int ns = mt.nsite;
  double F0 = 0, F1 = 0, F2 = 0, tx = 0, ty = 0, tz = 0;
  for (int s = 0; s < ns; s++) {
    const double r0 = t[0] * mt.bx[s] + t[1] * mt.by[s] + t[2] * mt.bz[s];
    const double r1 = t[3] * mt.bx[s] + t[4] * mt.by[s] + t[5] * mt.bz[s];
    const double r2 = t[6] * mt.bx[s] + t[7] * mt.by[s] + t[8] * mt.bz[s];
    const int k = i * ns + s;
    F0 = F0 + fx[k];
    F1 = F1 + fy[k];
    F2 = F2 + fz[k];
    tx = tx + r1 * fz[k] - r2 * fy[k];
    ty = ty + r2 * fx[k] - r0 * fz[k];
    tz = tz + r0 * fy[k] - r1 * fx[k];
  }
  force[i] = F0;
  force[n + i] = F1;
  force[2 * n + i] = F2;
  torque[i] = t[0] * tx + t[3] * ty + t[6] * tz;
  torque[n + i] = t[1] * tx + t[4] * ty + t[7] * tz;
  torque[2 * n + i] = t[2] * tx + t[5] * ty + t[8] * tz;
}

// CutOff_Dimmer2 (CutoffModule.F90:220-278) for a caller-given list, in the
// reference's operation order.
__global__ void k_cutoff(Geom g, int64_t np, const int *pi, const int *pj,
                         const double *com, const double *o, double *er,
                         double *fr) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int n = g.n, i = pi[k], j = pj[k];
  double d0 = com[i] - com[j], d1 = com[n + i] - com[n + j], d2 = com[2 * n + i] - com[2 * n + j];
  d0 = d0 - o[k];
  d1 = d1 - o[np + k];
  d2 = d2 - o[2 * np + k];
  const double dd = d0 * d0 + d1 * d1 + d2 * d2;
  double hh = 0.0, dh = 0.0;
  if (dd < g.rout * g.rout) {
    if (dd > g.rin * g.rin) {
      const double dr = sqrt(dd), u = dr - g.rout, l = dr - g.rin, uu = u * u, ll = l * l;
      hh = g.c1 * uu * u * (10.0 * ll - 5.0 * u * l + uu);
      dh = -(g.c2 * uu * ll / dr);
    } else {
      hh = 1.0;
    }
  }
  er[k] = hh;
  fr[k] = dh;
}

int cutoff_once(System &s, int64_t np, const int32_t *pi_h, const int32_t *pj_h,
                const double *com_h, const double *ox, const double *oy,
                const double *oz, double *er_h, double *fr_h) {
  if (np == 0) return 0;
  const int n = s.n;
  for (int64_t k = 0; k < np; k++)  // indices come from the caller: check them
    if (pi_h[k] < 0 || pi_h[k] >= n || pj_h[k] < 0 || pj_h[k] >= n) return G6_EARG;
  int *ij = nullptr;
  double *d = nullptr;
  G6_CHECK(hipMalloc(&ij, sizeof(int) * 2 * np));
  G6_CHECK(hipMalloc(&d, sizeof(double) * (3 * n + 5 * np)));
  double *com = d, *o = d + 3 * n, *er = o + 3 * np, *fr = er + np;
  G6_CHECK(hipMemcpyAsync(ij, pi_h, sizeof(int) * np, hipMemcpyHostToDevice, s.stream));
  G6_CHECK(hipMemcpyAsync(ij + np, pj_h, sizeof(int) * np, hipMemcpyHostToDevice, s.stream));
  G6_CHECK(hipMemcpyAsync(com, com_h, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s.stream));
  G6_CHECK(hipMemcpyAsync(o, ox, sizeof(double) * np, hipMemcpyHostToDevice, s.stream));
  G6_CHECK(hipMemcpyAsync(o + np, oy, sizeof(double) * np, hipMemcpyHostToDevice, s.stream));
  G6_CHECK(hipMemcpyAsync(o + 2 * np, oz, sizeof(double) * np, hipMemcpyHostToDevice, s.stream));
  k_cutoff<<<(unsigned)((np + 255) / 256), 256, 0, s.stream>>>(s.geo, np, ij, ij + np, com, o,
                                                               er, fr);
  G6_CHECK(hipGetLastError());
  G6_CHECK(hipMemcpyAsync(er_h, er, sizeof(double) * np, hipMemcpyDeviceToHost, s.stream));
  G6_CHECK(hipMemcpyAsync(fr_h, fr, sizeof(double) * np, hipMemcpyDeviceToHost, s.stream));
  G6_CHECK(hipStreamSynchronize(s.stream));
  G6_CHECK(hipFree(ij));
  G6_CHECK(hipFree(d));
  return 0;
}

int sites_once(System &s, const double *com_h, const double *q_h, double *x, double *y,
               double *z) {
  const int n = s.n, ns = s.ns;
  double *d = nullptr;
  G6_CHECK(hipMalloc(&d, sizeof(double) * (7 * n + 3 * n * ns)));
  double *com = d, *q = d + 3 * n, *o = d + 7 * n;
  G6_CHECK(hipMemcpyAsync(com, com_h, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s.stream));
  G6_CHECK(hipMemcpyAsync(q, q_h, sizeof(double) * 4 * n, hipMemcpyHostToDevice, s.stream));
  k_sites<<<(n + 255) / 256, 256, 0, s.stream>>>(s.mt, n, com, q, o, o + n * ns, o + 2 * n * ns);
  G6_CHECK(hipGetLastError());
  G6_CHECK(hipMemcpyAsync(x, o, sizeof(double) * n * ns, hipMemcpyDeviceToHost, s.stream));
  G6_CHECK(hipMemcpyAsync(y, o + n * ns, sizeof(double) * n * ns, hipMemcpyDeviceToHost, s.stream));
  G6_CHECK(hipMemcpyAsync(z, o + 2 * n * ns, sizeof(double) * n * ns, hipMemcpyDeviceToHost,
                          s.stream));
  G6_CHECK(hipStreamSynchronize(s.stream));
  G6_CHECK(hipFree(d));
  return 0;
}

int collect_once(System &s, const double *q_h, const double *fx, const double *fy,
                 const double *fz, double *force, double *torque) {
  const int n = s.n, ns = s.ns;
  double *d = nullptr;
  G6_CHECK(hipMalloc(&d, sizeof(double) * (10 * n + 3 * n * ns)));
  double *q = d, *F = d + 4 * n, *T = d + 7 * n, *f = d + 10 * n;
  G6_CHECK(hipMemcpyAsync(q, q_h, sizeof(double) * 4 * n, hipMemcpyHostToDevice, s.stream));
  G6_CHECK(hipMemcpyAsync(f, fx, sizeof(double) * n * ns, hipMemcpyHostToDevice, s.stream));
  G6_CHECK(hipMemcpyAsync(f + n * ns, fy, sizeof(double) * n * ns, hipMemcpyHostToDevice,
                          s.stream));
  G6_CHECK(hipMemcpyAsync(f + 2 * n * ns, fz, sizeof(double) * n * ns, hipMemcpyHostToDevice,
                          s.stream));
  k_collect<<<(n + 255) / 256, 256, 0, s.stream>>>(s.mt, n, q, f, f + n * ns, f + 2 * n * ns,
                                                    F, T);
  G6_CHECK(hipGetLastError());
  G6_CHECK(hipMemcpyAsync(force, F, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s.stream));
  G6_CHECK(hipMemcpyAsync(torque, T, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s.stream));
  G6_CHECK(hipStreamSynchronize(s.stream));
  G6_CHECK(hipFree(d));
  return 0;
}

}  // namespace g6
