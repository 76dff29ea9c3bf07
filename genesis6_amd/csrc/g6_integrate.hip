// Per-molecule kernels of the module-seam entry points:
//   rigid_setmatrix + rigid_setsiteposition0  RigidModule.F90:62-128 (k_sites)
//   rigid_collectforce    RigidModule.F90:275-317 (k_collect)
//   CutOff_Dimmer2        CutoffModule.F90:220-278 (k_cutoff)
// All FP64, evaluated in the reference's operation order (no contraction).
#include "g6_internal.h"

#pragma clang fp contract(off)

namespace g6 {

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
  char *base = nullptr;
  const int rs = seam_reserve(s, seam_piece(sizeof(int) * 2 * np) +
                                     seam_piece(sizeof(double) * (3 * n + 5 * np)),
                              &base);
  if (rs != 0) return rs;
  Carve cv{base};
  int *ij = cv.take<int>(2 * (size_t)np);
  double *d = cv.take<double>(3 * (size_t)n + 5 * (size_t)np);
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
  return 0;
}

int sites_once(System &s, const double *com_h, const double *q_h, double *x, double *y,
               double *z) {
  const int n = s.n, ns = s.ns;
  char *base = nullptr;
  const int rs = seam_reserve(s, seam_piece(sizeof(double) * (7 * n + 3 * n * ns)), &base);
  if (rs != 0) return rs;
  double *d = reinterpret_cast<double *>(base);
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
  return 0;
}

int collect_once(System &s, const double *q_h, const double *fx, const double *fy,
                 const double *fz, double *force, double *torque) {
  const int n = s.n, ns = s.ns;
  char *base = nullptr;
  const int rs = seam_reserve(s, seam_piece(sizeof(double) * (10 * n + 3 * n * ns)), &base);
  if (rs != 0) return rs;
  double *d = reinterpret_cast<double *>(base);
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
  return 0;
}

}  // namespace g6
