// The MD step in two kernels (the path g6_step runs):
//
//   k_md_pre   one thread per molecule: Nose_Predict (block 0), rigid_predict,
//              rigid_setmatrix + site offsets, Address_Assign binning into
//              fixed-capacity cell lists, and the Verlet-skin test that
//              decides whether the neighbour lists are rebuilt this step.
//   k_md_main  one wavefront per molecule i:
//              (1) if a rebuild is due, i's own neighbour row: all j != i in
//                  the 27 (or fewer, for ndiv < 3) surrounding cells with the
//                  reference's FP64 minimum-image test dd < (rc+skin)^2;
//              (2) the pair forces of i with every listed j whose FP64
//                  COM distance passes dd < rc^2 - the reference's
//                  Interaction_Compress2 test, so the interacting set is the
//                  reference's - via pair_sites (g6_pair.h);
//              (3) wave reduction of i's force and lab torque (the work of
//                  rigid_collectforce) and, on lane 0, rigid_correct for i;
//              (4) block partials of energy, virial and kinetic terms; the
//                  last block to finish reduces them in a fixed order and
//                  applies Nose_Correct (NoseModule.F90:43-72).
// Each unordered pair is evaluated in both rows, so no molecule's force
// needs an atomic and the FP64 results are deterministic; energies and the
// virial are halved accordingly.  Positions read by other waves come from
// com_p (the predicted COMs), never from the state being corrected.
#include "g6_pair.h"

namespace g6 {

constexpr double kG5_0 = 3.0 / 16.0, kG5_1 = 251.0 / 360.0, kG5_3 = 11.0 / 18.0,
                 kG5_4 = 1.0 / 6.0, kG5_5 = 1.0 / 60.0;
constexpr double kG4_0 = 251.0 / 720.0, kG4_2 = 11.0 / 12.0, kG4_3 = 1.0 / 3.0,
                 kG4_4 = 1.0 / 24.0;
#define G6_RGAS_MD (6.0225e23 * 1.380066e-23)

__device__ inline double wsum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__device__ inline void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- kernel A --------------------------------------------------------------
__global__ __launch_bounds__(64) void k_md_pre(ModelTable mt, Geom g, MdBuf b,
                                               int nose_active, int predict) {
#pragma clang fp contract(off)
  const int n = g.n;
  if (predict && nose_active && blockIdx.x == 0 && threadIdx.x == 0) {
    double *z = b.nose;
    z[3] = z[3] + z[4] + z[5] + z[6] + z[7];
    z[4] = z[4] + 2.0 * z[5] + 3.0 * z[6] + 4.0 * z[7];
    z[5] = z[5] + 3.0 * z[6] + 6.0 * z[7];
    z[6] = z[6] + 4.0 * z[7];
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double *com = b.com, *dx = b.dx, *w = b.w, *quat = b.quat;
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
  double p[3];
  for (int a = 0; a < 3; a++) {
    p[a] = com[a * n + i];
    b.com_p[a * n + i] = p[a];
  }
  const double qa = quat[i], qb = quat[n + i], qc = quat[2 * n + i], qd = quat[3 * n + i];
  double t[9];
  t[0] = (qa * qa + qb * qb - (qc * qc + qd * qd));
  t[1] = -2.0 * (qa * qd + qb * qc);
  t[2] = 2.0 * (qb * qd - qa * qc);
  t[3] = 2.0 * (qa * qd - qb * qc);
  t[4] = qa * qa + qc * qc - (qb * qb + qd * qd);
  t[5] = -2.0 * (qa * qb + qc * qd);
  t[6] = 2.0 * (qa * qc + qb * qd);
  t[7] = 2.0 * (qa * qb - qc * qd);
  t[8] = qa * qa + qd * qd - (qb * qb + qc * qc);
  for (int a = 0; a < 9; a++) b.tmat[a * n + i] = t[a];
  for (int s = 0; s < mt.nsite; s++) {
    const double r0 = t[0] * mt.bx[s] + t[1] * mt.by[s] + t[2] * mt.bz[s];
    const double r1 = t[3] * mt.bx[s] + t[4] * mt.by[s] + t[5] * mt.bz[s];
    const double r2 = t[6] * mt.bx[s] + t[7] * mt.by[s] + t[8] * mt.bz[s];
    b.intra[(s * 3 + 0) * n + i] = r0;
    b.intra[(s * 3 + 1) * n + i] = r1;
    b.intra[(s * 3 + 2) * n + i] = r2;
    b.intraf[(s * 3 + 0) * n + i] = (float)r0;
    b.intraf[(s * 3 + 1) * n + i] = (float)r1;
    b.intraf[(s * 3 + 2) * n + i] = (float)r2;
  }
  int jc[3];
  for (int a = 0; a < 3; a++) {
    double d = p[a];
    const double o = round(d * g.invL) * g.L;
    d = d - o;
    int j = (int)round(d * g.cellinv);
    if (j < 0) j += g.ndiv;
    j = ((j % g.ndiv) + g.ndiv) % g.ndiv;
    jc[a] = j;
  }
  const int c = (jc[2] * g.ndiv + jc[1]) * g.ndiv + jc[0];
  b.cell[i] = c;
  const int slot = atomicAdd(&b.cellcnt[c], 1);
  if (slot < b.cell_cap) b.cellres[(size_t)c * b.cell_cap + slot] = i;
  else b.flags[0] = 3;  // cell capacity exceeded
  if (b.half_skin2 > 0.0) {
    const double e0 = p[0] - b.com_ref[i], e1 = p[1] - b.com_ref[n + i],
                 e2 = p[2] - b.com_ref[2 * n + i];
    if (e0 * e0 + e1 * e1 + e2 * e2 > b.half_skin2) b.flags[1] = 1;
  }
}

// ---- kernel B --------------------------------------------------------------
__device__ inline void correct_one(const ModelTable &mt, int n, int i, double dt,
                                   double zeta0, const double *F, const double *tl,
                                   const double *tmat, double *com, double *dx, double *w,
                                   double *quat, double *ek) {
#pragma clang fp contract(off)
  double t[9];
  for (int a = 0; a < 9; a++) t[a] = tmat[a * n + i];
  const double tq[3] = {t[0] * tl[0] + t[3] * tl[1] + t[6] * tl[2],
                        t[1] * tl[0] + t[4] * tl[1] + t[7] * tl[2],
                        t[2] * tl[0] + t[5] * tl[1] + t[8] * tl[2]};
  const double nc = zeta0 * dt, ext = nc * 0.5, exr = nc;
  const double tf = dt * dt * 0.5 * (1.0 / mt.mass);
#define DX(s, a) dx[((s) * 3 + (a)) * n + i]
#define W(s, a) w[((s) * 3 + (a)) * n + i]
#define Q(s, a) quat[((s) * 4 + (a)) * n + i]
  for (int a = 0; a < 3; a++) {
    double nf = F[a] * tf;
    nf = nf - DX(0, a) * ext;
    const double d = nf - DX(1, a);
    com[a * n + i] = com[a * n + i] + kG5_0 * d;
    DX(0, a) = DX(0, a) + kG5_1 * d;
    DX(1, a) = nf;
    DX(2, a) = DX(2, a) + kG5_3 * d;
    DX(3, a) = DX(3, a) + kG5_4 * d;
    DX(4, a) = DX(4, a) + kG5_5 * d;
  }
  const double wx = W(0, 0), wy = W(0, 1), wz = W(0, 2);
  double qa = Q(0, 0), qb = Q(0, 1), qc = Q(0, 2), qd = Q(0, 3);
  const double nwa = (-wx * qb + wy * qc - wz * qd) * dt * 0.5;
  const double nwb = (wx * qa - wy * qd - wz * qc) * dt * 0.5;
  const double nwc = (-wx * qd - wy * qa + wz * qb) * dt * 0.5;
  const double nwd = (wx * qc + wy * qb + wz * qa) * dt * 0.5;
  const double dwa = nwa - Q(1, 0), dwb = nwb - Q(1, 1), dwc = nwc - Q(1, 2), dwd = nwd - Q(1, 3);
  Q(0, 0) = Q(0, 0) + kG4_0 * dwa;
  Q(0, 1) = Q(0, 1) + kG4_0 * dwb;
  Q(0, 2) = Q(0, 2) + kG4_0 * dwc;
  Q(0, 3) = Q(0, 3) + kG4_0 * dwd;
  double nf[3];
  nf[0] = (tq[0] + (mt.Iyy - mt.Izz) * wy * wz) * dt * (1.0 / mt.Ixx);
  nf[1] = (tq[1] + (mt.Izz - mt.Ixx) * wz * wx) * dt * (1.0 / mt.Iyy);
  nf[2] = (tq[2] + (mt.Ixx - mt.Iyy) * wx * wy) * dt * (1.0 / mt.Izz);
  for (int a = 0; a < 3; a++) {
    nf[a] = nf[a] - W(0, a) * exr;
    const double d = nf[a] - W(1, a);
    W(0, a) = W(0, a) + kG4_0 * d;
    W(1, a) = nf[a];
    W(2, a) = W(2, a) + kG4_2 * d;
    W(3, a) = W(3, a) + kG4_3 * d;
    W(4, a) = W(4, a) + kG4_4 * d;
  }
  const double w1x = W(0, 0), w1y = W(0, 1), w1z = W(0, 2);
  const double w2x = W(1, 0), w2y = W(1, 1), w2z = W(1, 2);
  const double w3x = W(2, 0), w3y = W(2, 1), w3z = W(2, 2);
  const double w4x = W(3, 0), w4y = W(3, 1), w4z = W(3, 2);
  qa = Q(0, 0); qb = Q(0, 1); qc = Q(0, 2); qd = Q(0, 3);
  const double pc = (-w1x * qd - w1y * qa + w1z * qb) * dt * 0.5;
  const double pb = (w1x * qa - w1y * qd - w1z * qc) * dt * 0.5;
  const double pd = (w1x * qc + w1y * qb + w1z * qa) * dt * 0.5;
  const double pa = (-w1x * qb + w1y * qc - w1z * qd) * dt * 0.5;
  Q(1, 0) = pa; Q(1, 1) = pb; Q(1, 2) = pc; Q(1, 3) = pd;
  const double rc_ = (-w2x * qd - w2y * qa + w2z * qb - w1x * pd - w1y * pa + w1z * pb) * dt * 0.25;
  const double rb = (w2x * qa - w2y * qd - w2z * qc + w1x * pa - w1y * pd - w1z * pc) * dt * 0.25;
  const double rd = (w2x * qc + w2y * qb + w2z * qa + w1x * pc + w1y * pb + w1z * pa) * dt * 0.25;
  const double ra = (-w2x * qb + w2y * qc - w2z * qd - w1x * pb + w1y * pc - w1z * pd) * dt * 0.25;
  Q(2, 0) = ra; Q(2, 1) = rb; Q(2, 2) = rc_; Q(2, 3) = rd;
  const double sc = (-w3x * qd - w3y * qa + w3z * qb - w1x * rd - w1y * ra + w1z * rb - w2x * pd -
                     w2y * pa + w2z * pb) * dt * (1.0 / 6.0);
  const double sb = (w3x * qa - w3y * qd - w3z * qc + w1x * ra - w1y * rd - w1z * rc_ + w2x * pa -
                     w2y * pd - w2z * pc) * dt * (1.0 / 6.0);
  const double sd = (w3x * qc + w3y * qb + w3z * qa + w1x * rc_ + w1y * rb + w1z * ra + w2x * pc +
                     w2y * pb + w2z * pa) * dt * (1.0 / 6.0);
  const double sa = (-w3x * qb + w3y * qc - w3z * qd - w1x * rb + w1y * rc_ - w1z * rd - w2x * pb +
                     w2y * pc - w2z * pd) * dt * (1.0 / 6.0);
  Q(3, 0) = sa; Q(3, 1) = sb; Q(3, 2) = sc; Q(3, 3) = sd;
  Q(4, 2) = (-w4x * qd - w4y * qa + w4z * qb - w3x * pd - w3y * pa + w3z * pb - w2x * rd -
             w2y * ra + w2z * rb - w1x * sd - w1y * sa + w1z * sb) * dt * (1.0 / 8.0);
  Q(4, 1) = (w4x * qa - w4y * qd - w4z * qc + w3x * pa - w3y * pd - w3z * pc + w2x * ra -
             w2y * rd - w2z * rc_ + w1x * sa - w1y * sd - w1z * sc) * dt * (1.0 / 8.0);
  Q(4, 3) = (w4x * qc + w4y * qb + w4z * qa + w3x * pc + w3y * pb + w3z * pa + w2x * rc_ +
             w2y * rb + w2z * ra + w1x * sc + w1y * sb + w1z * sa) * dt * (1.0 / 8.0);
  Q(4, 0) = (-w4x * qb + w4y * qc - w4z * qd - w3x * pb + w3y * pc - w3z * pd - w2x * rb +
             w2y * rc_ - w2z * rd - w1x * sb + w1y * sc - w1z * sd) * dt * (1.0 / 8.0);
  ek[0] = DX(0, 0) * DX(0, 0) + DX(0, 1) * DX(0, 1) + DX(0, 2) * DX(0, 2);
  ek[1] = W(0, 0) * W(0, 0);
  ek[2] = W(0, 1) * W(0, 1);
  ek[3] = W(0, 2) * W(0, 2);
#undef DX
#undef W
#undef Q
}

// neighbour-row build of molecule i into nbr[i*maxnb ...], in cell-stencil
// order; returns the row length (also stored).
__device__ inline int build_row(const Geom &g, const MdBuf &b, int i, int lane,
                                int wid, int (*s_beg)[32], int (*s_pre)[33]) {
  const int n = g.n;
  const int c = b.cell[i];
  const int nd = g.ndiv;
  const int cx = c % nd, cy = (c / nd) % nd, cz = c / (nd * nd);
  const int *cnt = b.cellcnt;
  const int *res = b.cellres;
  int len = 0;
  if (lane < b.nfull) {
    const int off = b.full[lane];
    const int ox = off & 0xff, oy = (off >> 8) & 0xff, oz = (off >> 16) & 0xff;
    const int c2 = (((cz + oz) % nd) * nd + (cy + oy) % nd) * nd + (cx + ox) % nd;
    s_beg[wid][lane] = c2 * b.cell_cap;
    len = min(cnt[c2], b.cell_cap);
  }
  int inc = len;
  for (int d = 1; d < 32; d <<= 1) {
    const int v = __shfl_up(inc, d, 64);
    if (lane >= d) inc += v;
  }
  if (lane < b.nfull) s_pre[wid][lane + 1] = inc;
  if (lane == 0) s_pre[wid][0] = 0;
  wsync();
  const int total = s_pre[wid][b.nfull];
  const double xi = b.com_p[i], yi = b.com_p[n + i], zi = b.com_p[2 * n + i];
  int base = 0;
  int *nj = b.nbr_j + (size_t)i * b.maxnb;
  int *nc = b.nbr_code + (size_t)i * b.maxnb;
  for (int t0 = 0; t0 < total; t0 += 64) {
    const int t = t0 + lane;
    bool pass = false;
    int j = 0, code = 0;
    if (t < total) {
      int s = 0;
      while (s + 1 < b.nfull && s_pre[wid][s + 1] <= t) s++;
      j = res[s_beg[wid][s] + (t - s_pre[wid][s])];
      if (j != i) {
        double d0 = xi - b.com_p[j], d1 = yi - b.com_p[n + j], d2 = zi - b.com_p[2 * n + j];
        const double k0 = round(d0 * g.invL), k1 = round(d1 * g.invL), k2 = round(d2 * g.invL);
        d0 = __dsub_rn(d0, __dmul_rn(k0, g.L));
        d1 = __dsub_rn(d1, __dmul_rn(k1, g.L));
        d2 = __dsub_rn(d2, __dmul_rn(k2, g.L));
        const double dd = __dadd_rn(__dadd_rn(__dmul_rn(d0, d0), __dmul_rn(d1, d1)),
                                    __dmul_rn(d2, d2));
        pass = dd < g.rr;
        const int f0 = (k0 == 0.0 && signbit(k0)) ? 0 : (int)k0 + 512;
        const int f1 = (k1 == 0.0 && signbit(k1)) ? 0 : (int)k1 + 512;
        const int f2 = (k2 == 0.0 && signbit(k2)) ? 0 : (int)k2 + 512;
        code = f0 | (f1 << 10) | (f2 << 20);
        if (pass && (fabs(k0) > 511.0 || fabs(k1) > 511.0 || fabs(k2) > 511.0)) b.flags[0] = 2;
      }
    }
    const unsigned long long mask = __ballot(pass);
    const int pos = base + __popcll(mask & ((1ull << lane) - 1ull));
    if (pass) {
      if (pos < b.maxnb) {
        nj[pos] = j;
        nc[pos] = code;
      } else {
        b.flags[0] = 4;  // neighbour-row capacity exceeded
      }
    }
    base += __popcll(mask);
  }
  wsync();
  const int m = min(base, b.maxnb);
  if (lane == 0) {
    b.nbr_cnt[i] = m;
    b.com_ref[i] = xi;
    b.com_ref[n + i] = yi;
    b.com_ref[2 * n + i] = zi;
  }
  return m;
}

template <int M, typename R>
__global__ __launch_bounds__(256) void k_md_main(ModelTable mt, Geom g, MdBuf b,
                                                 int nose_active, double dt, double *out) {
  using T = Traits<M>;
  __shared__ int s_beg[4][32], s_pre[4][33];
  __shared__ double s_red[4][16];
  __shared__ int s_last;
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = g.n;
  const int i = blockIdx.x * 4 + wid;
  const bool active = i < n;
  double acc[16];  // ep, vir[9], npair, ek[4], list pairs
#pragma unroll
  for (int q = 0; q < 16; q++) acc[q] = 0.0;
  if (active) {
    int m;
    if (b.flags[1]) {
      m = build_row(g, b, i, lane, wid, s_beg, s_pre);
      acc[15] = 0.5 * m;
    } else {
      m = b.nbr_cnt[i];
    }
    const R *intra = reinterpret_cast<const R *>(sizeof(R) == 4 ? (const void *)b.intraf
                                                                 : (const void *)b.intra);
    R ri[T::NS][3];
#pragma unroll
    for (int s = 0; s < T::NS; s++)
#pragma unroll
      for (int a = 0; a < 3; a++) ri[s][a] = intra[(s * 3 + a) * n + i];
    const double xi = b.com_p[i], yi = b.com_p[n + i], zi = b.com_p[2 * n + i];
    const int *nj = b.nbr_j + (size_t)i * b.maxnb;
    const int *nc = b.nbr_code + (size_t)i * b.maxnb;
    double F[3] = {0, 0, 0}, Ti[3] = {0, 0, 0}, ep = 0, v[9] = {0}, npi = 0;
    for (int k = lane; k < m; k += 64) {
      const int j = nj[k];
      const int code = nc[k];
      const int fa = code & 1023, fb = (code >> 10) & 1023, fc = (code >> 20) & 1023;
      const double ox = fa == 0 ? -0.0 : (double)(fa - 512) * g.L;
      const double oy = fb == 0 ? -0.0 : (double)(fb - 512) * g.L;
      const double oz = fc == 0 ? -0.0 : (double)(fc - 512) * g.L;
      const double d0 = __dsub_rn(__dsub_rn(xi, b.com_p[j]), ox);
      const double d1 = __dsub_rn(__dsub_rn(yi, b.com_p[n + j]), oy);
      const double d2 = __dsub_rn(__dsub_rn(zi, b.com_p[2 * n + j]), oz);
      const double dd64 = __dadd_rn(__dadd_rn(__dmul_rn(d0, d0), __dmul_rn(d1, d1)),
                                    __dmul_rn(d2, d2));
      if (!(dd64 < g.rc2)) continue;
      R rj[T::NS][3];
#pragma unroll
      for (int s = 0; s < T::NS; s++)
#pragma unroll
        for (int a = 0; a < 3; a++) rj[s][a] = intra[(s * 3 + a) * n + j];
      R hh, dh;
      cutoff_sw<R>(g, R(dd64), hh, dh);
      R fi[T::NS][3], fj[T::NS][3];
      const R e = pair_sites<M, R>(mt, hh, dh, R(d0), R(d1), R(d2), ri, rj, fi, fj);
      R Fx = 0, Fy = 0, Fz = 0, tix = 0, tiy = 0, tiz = 0;
#pragma unroll
      for (int s = 0; s < T::NS; s++) {
        Fx += fi[s][0]; Fy += fi[s][1]; Fz += fi[s][2];
        tix += ri[s][1] * fi[s][2] - ri[s][2] * fi[s][1];
        tiy += ri[s][2] * fi[s][0] - ri[s][0] * fi[s][2];
        tiz += ri[s][0] * fi[s][1] - ri[s][1] * fi[s][0];
      }
      F[0] += (double)Fx; F[1] += (double)Fy; F[2] += (double)Fz;
      Ti[0] += (double)tix; Ti[1] += (double)tiy; Ti[2] += (double)tiz;
      ep += (double)e;
      npi += 1.0;
      v[0] += d0 * (double)Fx; v[1] += d0 * (double)Fy; v[2] += d0 * (double)Fz;
      v[3] += d1 * (double)Fx; v[4] += d1 * (double)Fy; v[5] += d1 * (double)Fz;
      v[6] += d2 * (double)Fx; v[7] += d2 * (double)Fy; v[8] += d2 * (double)Fz;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
      F[a] = wsum(F[a]);
      Ti[a] = wsum(Ti[a]);
    }
    acc[0] = 0.5 * wsum(ep);
#pragma unroll
    for (int a = 0; a < 9; a++) acc[1 + a] = 0.5 * wsum(v[a]);
    acc[10] = 0.5 * wsum(npi);
    if (lane == 0) {
      if (out) {
        // eval mode: molecule force (lab) and body torque, state untouched
        double t[9];
        for (int a = 0; a < 9; a++) t[a] = b.tmat[a * n + i];
        out[i] = F[0];
        out[n + i] = F[1];
        out[2 * n + i] = F[2];
        out[3 * n + i] = t[0] * Ti[0] + t[3] * Ti[1] + t[6] * Ti[2];
        out[4 * n + i] = t[1] * Ti[0] + t[4] * Ti[1] + t[7] * Ti[2];
        out[5 * n + i] = t[2] * Ti[0] + t[5] * Ti[1] + t[8] * Ti[2];
      } else {
        double ek[4];
        correct_one(mt, n, i, dt, nose_active ? b.nose[3] : 0.0, F, Ti, b.tmat, b.com, b.dx,
                    b.w, b.quat, ek);
        for (int q = 0; q < 4; q++) acc[11 + q] = ek[q];
      }
    }
  }
  // block partials: lane 0 of each wave holds the wave's values
  if (lane == 0)
    for (int q = 0; q < 16; q++) s_red[wid][q] = acc[q];
  __syncthreads();
  if (threadIdx.x < 16) {
    const int q = threadIdx.x;
    b.blockpart[(size_t)blockIdx.x * 16 + q] = s_red[0][q] + s_red[1][q] + s_red[2][q] + s_red[3][q];
  }
  // last block to finish closes the step
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) s_last = (atomicAdd(b.done, 1) == (int)gridDim.x - 1);
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  __shared__ double red[16][256];
  double a2[16];
  for (int q = 0; q < 16; q++) a2[q] = 0.0;
  for (int k = threadIdx.x; k < (int)gridDim.x; k += 256)
    for (int q = 0; q < 16; q++) a2[q] += b.blockpart[(size_t)k * 16 + q];
  for (int q = 0; q < 16; q++) red[q][threadIdx.x] = a2[q];
  // the cell lists are consumed: clear the counts for the next step
  for (int c = threadIdx.x; c <= g.ncell; c += 256) b.cellcnt[c] = 0;
  __syncthreads();
  for (int d = 128; d >= 1; d >>= 1) {
    if ((int)threadIdx.x < d)
      for (int q = 0; q < 16; q++) red[q][threadIdx.x] += red[q][threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma clang fp contract(off)
    double *th = b.thermo;
    const double ep = red[0][0];
    const double ekt = red[11][0] * mt.mass * 0.5 * (1.0 / (dt * dt));
    const double ekr = (red[12][0] * mt.Ixx + red[13][0] * mt.Iyy + red[14][0] * mt.Izz) * 0.5;
    const double dof = 6.0 * n;
    const double temp = 10.0 * 2.0 * (ekt + ekr) / (G6_RGAS_MD * dof);
    double h = (ep + ekt + ekr) * 10.0;
    if (nose_active && !out) {
      double *z = b.nose;
      const double zetaf = (temp - z[0]) / z[1] - z[2] * z[3];
      const double dz2 = zetaf - z[4];
      z[3] = z[3] + kG4_0 * dz2;
      z[4] = zetaf;
      z[5] = z[5] + kG4_2 * dz2;
      z[6] = z[6] + kG4_3 * dz2;
      z[7] = z[7] + kG4_4 * dz2;
      z[8] = z[8] + z[3];
      h = h + (dof * z[0] * (z[8] - 0.5 * z[3]) * dt) * G6_RGAS_MD +
          (z[3] * z[3] * z[1] * dt * dof * 0.5) * G6_RGAS_MD;
    }
    th[0] = ep;
    th[1] = ekt;
    th[2] = ekr;
    th[3] = temp;
    for (int q = 0; q < 9; q++) th[4 + q] = red[1 + q][0];
    th[13] = h;
    th[14] = red[10][0];
    if (b.flags[1]) th[15] = red[15][0];
    if (b.half_skin2 > 0.0) b.flags[1] = 0;
    *b.done = 0;
  }
}

void md_launch_pre(System &s, bool predict) {
  k_md_pre<<<(s.n + 63) / 64, 64, 0, s.stream>>>(s.mt, s.geo_md, s.md, s.nose_active ? 1 : 0,
                                                  predict ? 1 : 0);
}

template <int M, typename R>
static void md_main_t(System &s, double *out) {
  k_md_main<M, R><<<s.nblk_md, 256, 0, s.stream>>>(s.mt, s.geo_md, s.md, s.nose_active ? 1 : 0,
                                                   s.dt, out);
}

void md_launch_main(System &s, double *out) {
  if (s.prec == G6_FP32) {
    if (s.model == G6_TIP4P) md_main_t<G6_TIP4P, float>(s, out);
    else if (s.model == G6_TIP5P) md_main_t<G6_TIP5P, float>(s, out);
    else md_main_t<G6_ST2, float>(s, out);
  } else {
    if (s.model == G6_TIP4P) md_main_t<G6_TIP4P, double>(s, out);
    else if (s.model == G6_TIP5P) md_main_t<G6_TIP5P, double>(s, out);
    else md_main_t<G6_ST2, double>(s, out);
  }
}

}  // namespace g6
