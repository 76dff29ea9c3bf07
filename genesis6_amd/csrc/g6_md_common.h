// Device helpers shared by the MD kernels (g6_md.hip: rows of both
// directions; g6_tile.hip: cell tiles evaluating each pair once): DPP wave
// sums, compensated sums, the box of the step, the per-molecule record
// layout and the step close.  Everything here is FP64 bookkeeping in the
// reference's operation order: no FMA contraction.
#pragma once
#include "g6_pair.h"

#pragma clang fp contract(off)

namespace g6 {

// The step's error code: the first one raised stays (a later symptom of the
// same failure, e.g. a runaway molecule's image shift, does not replace it).
__device__ inline void raise_flag(int *flags, int code) { atomicCAS(flags, 0, code); }


constexpr double kG5_0 = 3.0 / 16.0, kG5_1 = 251.0 / 360.0, kG5_3 = 11.0 / 18.0,
                 kG5_4 = 1.0 / 6.0, kG5_5 = 1.0 / 60.0;
constexpr double kG4_0 = 251.0 / 720.0, kG4_2 = 11.0 / 12.0, kG4_3 = 1.0 / 3.0,
                 kG4_4 = 1.0 / 24.0;
#define G6_RGAS_MD (6.0225e23 * 1.380066e-23)

// DPP move of a double across lanes (both halves); every row written.
template <int CTRL>
__device__ inline double dpp_d(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double,
                            ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
}

// As dpp_d for the rows selected by RM; the other rows read 0.
template <int CTRL, int RM>
__device__ inline double dpp_dm(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, RM, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, RM, 0xF, false);
  return __builtin_bit_cast(double,
                            ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
}

// Sum over the 64 lanes of a wave in a fixed order; the total is in lane 63.
// Steps: xor 1, xor 2 (quad_perm), half-row mirror, row mirror, then
// row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3.
// Every lane of the wave must be active.
__device__ inline double wsum63(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  v += dpp_dm<0x142, 0xA>(v);
  v += dpp_dm<0x143, 0xC>(v);
  return v;
}

// Compensated sums (Knuth's TwoSum): the pair (s, c) holds s + c, the
// rounding error of every addition carried in c.
__device__ inline void two_sum_acc(double &s, double &c, double x) {
  const double t = s + x;
  const double bp = t - s;
  c = c + ((s - (t - bp)) + (x - bp));
  s = t;
}
__device__ inline void dd_add(double &s, double &c, double so, double co) {
  const double t = s + so;
  const double bp = t - s;
  const double e = (s - (t - bp)) + (so - bp);
  c = (c + co) + e;
  s = t;
}
// rowsum16 on compensated pairs (the same lane pattern)
__device__ inline void rowsum16_dd(double &s, double &c) {
  dd_add(s, c, dpp_d<0xB1>(s), dpp_d<0xB1>(c));
  dd_add(s, c, dpp_d<0x4E>(s), dpp_d<0x4E>(c));
  dd_add(s, c, dpp_d<0x141>(s), dpp_d<0x141>(c));
  dd_add(s, c, dpp_d<0x140>(s), dpp_d<0x140>(c));
}

__device__ inline void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// rigid_predict of quaternion component A (levels 0 and 1)
template <int A>
__device__ inline void predict_qc(double (&S)[20]) {
  S[A] = S[A] + S[4 + A] + S[8 + A] + S[12 + A] + S[16 + A];
  S[4 + A] = S[4 + A] + 2.0 * S[8 + A] + 3.0 * S[12 + A] + 4.0 * S[16 + A];
}

// components C0 and C0+1 of the five quaternion levels back to HBM
template <int C0>
__device__ inline void store_qc(const MdBuf &b, int n, int i, const double (&S)[20]) {
#pragma unroll
  for (int s = 0; s < 5; s++) {
    b.quat[(s * 4 + C0) * n + i] = S[s * 4 + C0];
    b.quat[(s * 4 + C0 + 1) * n + i] = S[s * 4 + C0 + 1];
  }
}

// The box of this step: the fixed box of the grid, or the Andersen
// barostat's current box in device memory (written by the step close).
struct BoxNow {
  double L[3], invL[3], cellinv[3];
};
// The kernels are instantiated per kind of box: a fixed cube (one edge for
// every axis, so the edges take as few registers as a scalar box), a fixed
// orthorhombic box, and the barostat's box in device memory.
constexpr int kBoxCube = 0, kBoxOrtho = 1, kBoxBaro = 2;
template <int BOX>
__device__ inline BoxNow box_now(const Geom &g, const MdBuf &b) {
  BoxNow r;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int e = BOX == kBoxCube ? 0 : a;
    r.L[a] = BOX == kBoxBaro ? b.box[a] : g.L[e];
    r.invL[a] = BOX == kBoxBaro ? b.box[3 + a] : g.invL[e];
    r.cellinv[a] = BOX == kBoxBaro ? b.box[6 + a] : g.cellinv[e];
  }
  return r;
}
// the MD grid's cells per axis (one count for a cube)
template <int BOX>
__device__ inline void grid_nd(const Geom &g, int (&nd)[3]) {
#pragma unroll
  for (int a = 0; a < 3; a++) nd[a] = g.nd[BOX == kBoxCube ? 0 : a];
}

// Andersen_Rescale (AndersenModule.F90:65-82, BoxModule.F90:167-183): in
// the orthorhombic mode every edge scales by (v0 / V)^(1/3), in the orz mode
// the z edge alone by v0 / V; then the inverses and the MD grid's 1/cell.
__device__ inline void andersen_rescale(const double *a, double *box) {
  const double vol = box_volume(box);
  if (box[16] == 2.0) {
    box[2] = box[2] * (a[2] / vol);
  } else {
    const double r2 = pow(a[2] / vol, 1.0 / 3.0);
    box[0] = box[0] * r2;
    box[1] = box[1] * r2;
    box[2] = box[2] * r2;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    box[3 + k] = 1.0 / box[k];
    box[6 + k] = 1.0 / (box[k] / box[12 + k]);
  }
}

// ---- step close: one block of 256 threads ---------------------------------
// mode bits: 1 the kinetic partials of k_md_pre + the pair partials of the
// previous k_md_main (bpart) form the thermo record, then Nose_Correct;
// 2 Nose_Predict; 4 clear the cell counts and rebuild flag of parity 1-p
// (8: both parities); 16 force evaluation: the pair partials alone go to the
// thermo record.  Runs as its own kernel (k_md_close) or at the start of
// block 0 of k_md_main, which neither reads nor writes what it touches.
template <bool NPT>
__device__ inline void close_body(const ModelTable &mt, const MdBuf &b, int n, double dt,
                                  int nblk_pre, const double *bpart, int nblk_main, int ncell,
                                  int nose_active, int mode, int parity, double box_vol,
                                  double (&s_red)[4][kRed]) {
  constexpr int NR = 12 + kEkLen;
  const int wid = threadIdx.x >> 6;
  // the work runs on threads 0..255; wider blocks (the tile kernel's) only
  // take part in the barriers
  if (threadIdx.x < 256) {
  // the first kinetic partials are fetched together with the pair partials
  // (one memory round trip fewer; the sums keep their order)
  double ek0[kEkLen];
  {
    const bool first = (mode & 1) && (int)threadIdx.x < nblk_pre;
#pragma unroll
    for (int q = 0; q < kEkLen; q++)
      ek0[q] = first ? b.blockek[(size_t)threadIdx.x * kEkLen + q] : 0.0;
  }
  {
    // the pair partials of k_md_main (each rounded once from its compensated
    // block sum), summed over blocks in a fixed order
    double ps[12];
#pragma unroll
    for (int q = 0; q < 12; q++) ps[q] = 0.0;
    if (mode & 17) {
#pragma unroll 4
      for (int k = threadIdx.x; k < nblk_main; k += 256)
#pragma unroll
        for (int q = 0; q < 12; q++) ps[q] += bpart[(size_t)k * kBP + q];
    }
#pragma unroll
    for (int q = 0; q < 12; q++) ps[q] = wsum63(ps[q]);
    if ((threadIdx.x & 63) == 63)
#pragma unroll
      for (int q = 0; q < 12; q++) s_red[wid][q] = ps[q];
  }
  {
    // the kinetic partials of k_md_pre
    double ek[kEkLen];
#pragma unroll
    for (int q = 0; q < kEkLen; q++) ek[q] = 0.0 + ek0[q];
    if (mode & 1) {
#pragma unroll 1
      for (int k = threadIdx.x + 256; k < nblk_pre; k += 256)
#pragma unroll
        for (int q = 0; q < kEkLen; q++) ek[q] += b.blockek[(size_t)k * kEkLen + q];
    }
#pragma unroll
    for (int q = 0; q < kEkLen; q++) ek[q] = wsum63(ek[q]);
    if ((threadIdx.x & 63) == 63)
#pragma unroll
      for (int q = 0; q < kEkLen; q++) s_red[wid][12 + q] = ek[q];
  }
  if (mode & 12) {
    const int p0 = (mode & 8) ? 0 : 1 - parity, p1 = (mode & 8) ? 1 : 1 - parity;
    for (int p = p0; p <= p1; p++) {
      int *cr = b.cellres + (size_t)p * ncell * b.cell_cap;
      for (int c = threadIdx.x; c < ncell; c += 256) cr[(size_t)c * b.cell_cap] = 0;
      if (threadIdx.x == 0 && b.half_skin2 > 0.0) b.flags[2 + p] = 0;
    }
  }
  }
  __syncthreads();
  // totals over the four waves, one quantity per thread, into row 0
  if (threadIdx.x < NR) {
    const int q = threadIdx.x;
    s_red[0][q] = s_red[0][q] + s_red[1][q] + s_red[2][q] + s_red[3][q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double *r = s_red[0];
    double *z = b.nose, *th = b.thermo;
    if (mode & 17) {
      th[0] = r[0];
      for (int q = 0; q < 9; q++) th[4 + q] = r[1 + q];
      th[14] = r[10];
      if (r[11] > 0) th[15] = r[11];
    }
    if (mode & 1) {
      const double ep = r[0];
      const double ekt = r[12] * mt.mass * 0.5 * (1.0 / (dt * dt));
      const double ekr = (r[13] * mt.Ixx + r[14] * mt.Iyy + r[15] * mt.Izz) * 0.5;
      const double dof = 6.0 * n;
      const double temp = 10.0 * 2.0 * (ekt + ekr) / (G6_RGAS_MD * dof);
      double h = (ep + ekt + ekr) * 10.0;
      double nose_ep = 0.0, nose_ek = 0.0;
      if (nose_active) {
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
      // kinetic tensor (RigidModule.F90:617: sums * mass * (1/dt^2)) and
      // pressure (Mixture.F90:2236-2255) in the box of this step
      const double kd = 1.0 / (dt * dt);
      const double kt[9] = {r[16] * mt.mass * kd, r[17] * mt.mass * kd, r[18] * mt.mass * kd,
                            r[17] * mt.mass * kd, r[19] * mt.mass * kd, r[20] * mt.mass * kd,
                            r[18] * mt.mass * kd, r[20] * mt.mass * kd, r[21] * mt.mass * kd};
      for (int q = 0; q < 9; q++) th[16 + q] = kt[q];
      const double vp = (kt[0] + th[4]) + (kt[4] + th[8]) + (kt[8] + th[12]);
      const double vol = NPT ? box_volume(b.box) : box_vol;
      const double press =
          (vp / 3.0) * 10.0 * (1.0 / 101326.0) / (vol * 6.0225e23 * (1e-10 * 1e-10 * 1e-10));
      th[25] = press;
      double and_ep = 0.0, and_ek = 0.0, volc = vol;
      if (NPT) {
        // Andersen_Correct (AndersenModule.F90:95-110) with this pressure
        double *a = b.anders;
        const double volp2 = (press - a[0]) / a[1] * 0.5;
        const double cvol = volp2 - a[4];
        a[2] = a[2] + kG5_0 * cvol;
        a[3] = a[3] + kG5_1 * cvol;
        a[4] = volp2;
        a[5] = a[5] + kG5_3 * cvol;
        a[6] = a[6] + kG5_4 * cvol;
        a[7] = a[7] + kG5_5 * cvol;
        andersen_rescale(a, b.box);
        volc = box_volume(b.box);
        // barostat energies of the log line (Mixture.F90:2436-2439), J/mol
        and_ep = a[0] * a[2] * 101326.0 * 6.0225e23 * (1e-10 * 1e-10 * 1e-10);
        and_ek = a[3] * a[3] * a[1] * 0.5 * 101326.0 * 6.0225e23 * (1e-10 * 1e-10 * 1e-10);
        th[13] = h + and_ep + and_ek;
      }
      th[26] = volc;
      th[27] = nose_ep;
      th[28] = nose_ek;
      th[29] = nose_active ? z[0] : 0.0;
      th[30] = and_ep;
      th[31] = and_ek;
    }
    if ((mode & 2) && nose_active) {
      z[3] = z[3] + z[4] + z[5] + z[6] + z[7];
      z[4] = z[4] + 2.0 * z[5] + 3.0 * z[6] + 4.0 * z[7];
      z[5] = z[5] + 3.0 * z[6] + 6.0 * z[7];
      z[6] = z[6] + 4.0 * z[7];
    }
    if (NPT && (mode & 2)) {
      // Andersen_Predict (AndersenModule.F90:84-93)
      double *a = b.anders;
      a[2] = a[2] + a[3] + a[4] + a[5] + a[6] + a[7];
      a[3] = a[3] + 2.0 * a[4] + 3.0 * a[5] + 4.0 * a[6] + 5.0 * a[7];
      a[4] = a[4] + 3.0 * a[5] + 6.0 * a[6] + 10.0 * a[7];
      a[5] = a[5] + 4.0 * a[6] + 10.0 * a[7];
      a[6] = a[6] + 5.0 * a[7];
      andersen_rescale(a, b.box);
    }
  }
}


// site offsets of the sites before the CoM site from a record
template <int M, typename R>
struct RecSites;
template <int M>
struct RecSites<M, float> {
  static constexpr int NV = (Traits<M>::COM * 3 + 3) / 4;  // float4 loads
  __device__ static void load(const double *rec, float (&r)[Traits<M>::NS][3]) {
    const float4 *p = reinterpret_cast<const float4 *>(rec + 4);
    float f[NV * 4];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const float4 q = p[v];
      f[4 * v] = q.x; f[4 * v + 1] = q.y; f[4 * v + 2] = q.z; f[4 * v + 3] = q.w;
    }
#pragma unroll
    for (int s = 0; s < Traits<M>::COM; s++)
#pragma unroll
      for (int a = 0; a < 3; a++) r[s][a] = f[s * 3 + a];
#pragma unroll
    for (int a = 0; a < 3; a++) r[Traits<M>::COM][a] = 0.0f;
  }
};
template <int M>
struct RecSites<M, double> {
  static constexpr int NV = (Traits<M>::COM * 3 + 1) / 2;  // double2 loads
  __device__ static void load(const double *rec, double (&r)[Traits<M>::NS][3]) {
    const double2 *p = reinterpret_cast<const double2 *>(rec + 4);
    double f[NV * 2];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const double2 q = p[v];
      f[2 * v] = q.x; f[2 * v + 1] = q.y;
    }
#pragma unroll
    for (int s = 0; s < Traits<M>::COM; s++)
#pragma unroll
      for (int a = 0; a < 3; a++) r[s][a] = f[s * 3 + a];
#pragma unroll
    for (int a = 0; a < 3; a++) r[Traits<M>::COM][a] = 0.0;
  }
};

// Sum over the 16 lanes of a DPP row in a fixed order; every lane of the
// row receives the row total.
__device__ inline double rowsum16(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  return v;
}

}  // namespace g6
