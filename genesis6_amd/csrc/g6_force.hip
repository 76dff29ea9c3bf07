// Force kernels of the module-seam entry points (the MD path is g6_md.hip).
//
// k_force_pairs / k_energy_pairs: the module-seam entry points force() and
// PairPotential() on caller-given site coordinates and ratios (one thread
// per pair, site forces scattered with atomics).
#include "g6_pair.h"

namespace g6 {

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// ---- module seam: force() / PairPotential() on host-given arrays ---------
template <int M, typename R>
__global__ __launch_bounds__(256) void k_force_pairs(
    ModelTable mt, int64_t np, const int *pi, const int *pj, const double *ox,
    const double *oy, const double *oz, const double *er, const double *fr,
    const double *x, const double *y, const double *z, double *fx, double *fy,
    double *fz, double *ep, double *vir, int energy_only) {
  using T = Traits<M>;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[9] = {0};
  if (k < np) {
    const int ii = pi[k] * T::NS, jj = pj[k] * T::NS;
    const int c = T::COM;
    double d0 = x[ii + c], d1 = y[ii + c], d2 = z[ii + c];
    d0 = d0 - x[jj + c];
    d1 = d1 - y[jj + c];
    d2 = d2 - z[jj + c];
    d0 = d0 - ox[k];
    d1 = d1 - oy[k];
    d2 = d2 - oz[k];
    R ri[T::NS][3], rj[T::NS][3];
#pragma unroll
    for (int s = 0; s < T::NS; s++) {
      ri[s][0] = R(x[ii + s] - x[ii + c]);
      ri[s][1] = R(y[ii + s] - y[ii + c]);
      ri[s][2] = R(z[ii + s] - z[ii + c]);
      rj[s][0] = R(x[jj + s] - x[jj + c]);
      rj[s][1] = R(y[jj + s] - y[jj + c]);
      rj[s][2] = R(z[jj + s] - z[jj + c]);
    }
    if (energy_only) {
      ep[k] = (double)pair_energy<M, R>(mt, R(er[k]), R(d0), R(d1), R(d2), ri, rj);
    } else {
      R fi[T::NS][3], fj[T::NS][3];
      ep[k] = (double)pair_sites<M, R>(mt, R(er[k]), R(fr[k]), R(d0), R(d1), R(d2), ri,
                                       rj, fi, fj);
      double Fs[3] = {0, 0, 0};
#pragma unroll
      for (int s = 0; s < T::NS; s++) {
        unsafeAtomicAdd(&fx[ii + s], (double)fi[s][0]);
        unsafeAtomicAdd(&fy[ii + s], (double)fi[s][1]);
        unsafeAtomicAdd(&fz[ii + s], (double)fi[s][2]);
        unsafeAtomicAdd(&fx[jj + s], -(double)fj[s][0]);
        unsafeAtomicAdd(&fy[jj + s], -(double)fj[s][1]);
        unsafeAtomicAdd(&fz[jj + s], -(double)fj[s][2]);
        Fs[0] += (double)fi[s][0]; Fs[1] += (double)fi[s][1]; Fs[2] += (double)fi[s][2];
      }
      const double dv[3] = {d0, d1, d2};
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) v[3 * a + b] = dv[a] * Fs[b];
    }
  }
  if (!energy_only) {
    __shared__ double red[4][9];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 9; a++) {
      const double t = wave_sum(v[a]);
      if (lane == 0) red[wid][a] = t;
    }
    __syncthreads();
    if (threadIdx.x < 9)
      unsafeAtomicAdd(&vir[threadIdx.x], red[0][threadIdx.x] + red[1][threadIdx.x] +
                                             red[2][threadIdx.x] + red[3][threadIdx.x]);
  }
}

template <int M, typename R>
static void launch_pairs_t(System &s, int64_t np, double **d, double *vir,
                           const int *pi, const int *pj, int eo) {
  const int nb = (int)((np + 255) / 256);
  k_force_pairs<M, R><<<nb, 256, 0, s.stream>>>(s.mt, np, pi, pj, d[0], d[1], d[2], d[3],
                                                 d[4], d[5], d[6], d[7], d[8], d[9], d[10],
                                                 d[11], vir, eo);
}

int force_pairs(System &s, int64_t np, const int32_t *pi_h, const int32_t *pj_h,
                const double *ox, const double *oy, const double *oz,
                const double *er, const double *fr, const double *x,
                const double *y, const double *z, double *fx, double *fy,
                double *fz, double *ep, double *vrmat, bool energy_only) {
  if (np < 0) return G6_EARG;
  for (int64_t k = 0; k < np; k++)  // indices come from the caller: check them
    if (pi_h[k] < 0 || pi_h[k] >= s.n || pj_h[k] < 0 || pj_h[k] >= s.n) return G6_EARG;
  const int nsites = s.n * s.ns;
  const size_t pb = sizeof(double) * (size_t)(np > 0 ? np : 1);
  const size_t sb = sizeof(double) * (size_t)nsites;
  const size_t ib = sizeof(int) * (size_t)(np > 0 ? np : 1);
  char *base = nullptr;
  const int rs = seam_reserve(s, 2 * seam_piece(ib) + 6 * seam_piece(pb) + 6 * seam_piece(sb) +
                                     seam_piece(sizeof(double) * 9),
                              &base);
  if (rs != 0) return rs;
  Carve cv{base};
  int *pi = cv.take<int>(np > 0 ? np : 1), *pj = cv.take<int>(np > 0 ? np : 1);
  double *d[12];
  for (int a = 0; a < 5; a++) d[a] = cv.take<double>(pb / sizeof(double));
  for (int a = 5; a < 11; a++) d[a] = cv.take<double>(sb / sizeof(double));
  d[11] = cv.take<double>(pb / sizeof(double));
  double *vir = cv.take<double>(9);
  hipStream_t st = s.stream;
  if (np > 0) {
    G6_CHECK(hipMemcpyAsync(pi, pi_h, sizeof(int) * np, hipMemcpyHostToDevice, st));
    G6_CHECK(hipMemcpyAsync(pj, pj_h, sizeof(int) * np, hipMemcpyHostToDevice, st));
    const double *src[5] = {ox, oy, oz, er, fr};
    for (int a = 0; a < 5; a++) {
      if (energy_only && a == 4) continue;
      G6_CHECK(hipMemcpyAsync(d[a], src[a], pb, hipMemcpyHostToDevice, st));
    }
  }
  const double *ssrc[3] = {x, y, z};
  for (int a = 0; a < 3; a++)
    G6_CHECK(hipMemcpyAsync(d[5 + a], ssrc[a], sb, hipMemcpyHostToDevice, st));
  if (!energy_only) {
    const double *fsrc[3] = {fx, fy, fz};
    for (int a = 0; a < 3; a++)
      G6_CHECK(hipMemcpyAsync(d[8 + a], fsrc[a], sb, hipMemcpyHostToDevice, st));
  }
  G6_CHECK(hipMemsetAsync(vir, 0, sizeof(double) * 9, st));
  if (np > 0) {
    const int eo = energy_only ? 1 : 0;
    if (s.prec == G6_FP32) {
      if (s.kmodel == G6_TIP4P) launch_pairs_t<G6_TIP4P, float>(s, np, d, vir, pi, pj, eo);
      else if (s.kmodel == G6_TIP5P) launch_pairs_t<G6_TIP5P, float>(s, np, d, vir, pi, pj, eo);
      else if (s.kmodel == G6_SPCE) launch_pairs_t<G6_SPCE, float>(s, np, d, vir, pi, pj, eo);
      else if (s.kmodel == G6_NVDE) launch_pairs_t<G6_NVDE, float>(s, np, d, vir, pi, pj, eo);
      else if (s.kmodel == G6_OPLSMEOH)
        launch_pairs_t<G6_OPLSMEOH, float>(s, np, d, vir, pi, pj, eo);
      else launch_pairs_t<G6_ST2, float>(s, np, d, vir, pi, pj, eo);
    } else {
      if (s.kmodel == G6_TIP4P) launch_pairs_t<G6_TIP4P, double>(s, np, d, vir, pi, pj, eo);
      else if (s.kmodel == G6_TIP5P) launch_pairs_t<G6_TIP5P, double>(s, np, d, vir, pi, pj, eo);
      else if (s.kmodel == G6_SPCE) launch_pairs_t<G6_SPCE, double>(s, np, d, vir, pi, pj, eo);
      else if (s.kmodel == G6_NVDE) launch_pairs_t<G6_NVDE, double>(s, np, d, vir, pi, pj, eo);
      else if (s.kmodel == G6_OPLSMEOH)
        launch_pairs_t<G6_OPLSMEOH, double>(s, np, d, vir, pi, pj, eo);
      else launch_pairs_t<G6_ST2, double>(s, np, d, vir, pi, pj, eo);
    }
    G6_CHECK(hipGetLastError());
    G6_CHECK(hipMemcpyAsync(ep, d[11], sizeof(double) * np, hipMemcpyDeviceToHost, st));
  }
  if (!energy_only) {
    double *fdst[3] = {fx, fy, fz};
    for (int a = 0; a < 3; a++)
      G6_CHECK(hipMemcpyAsync(fdst[a], d[8 + a], sb, hipMemcpyDeviceToHost, st));
    G6_CHECK(hipMemcpyAsync(vrmat, vir, sizeof(double) * 9, hipMemcpyDeviceToHost, st));
  }
  G6_CHECK(hipStreamSynchronize(st));
  return 0;
}

}  // namespace g6
