// Neighbour arrays of the order parameters (Neighbor.F90): neighbors_all2
// (91-159), neighbor1_compress2 (195-215) and neighbors_displacement
// (401-457), all built on the insertion of Neighbor.F90:20-88.
//
// The reference inserts entries one at a time: appended (ORDER_NONE), or in
// front of the first entry with a strictly larger key (distance or label),
// cut to an optional limit.  Equal keys therefore keep their insertion order,
// and an entry that falls past the limit can never move back in front of it,
// so each array is its candidates stably sorted by (key, insertion order)
// and cut to the limit.  Here one wave owns one array: it stages the
// candidates in LDS and writes each to its rank under that total key.  For
// neighbors_all2 the candidates are binned first (count, offsets, atomic
// slots); the insertion order of a molecule's candidates is the pair index,
// so the output does not depend on the order the atomics filled the bins.
//
// Layout: a table of arrays is nnear[n] plus near, dx, dy, dz, dd of n rows
// of kMaxNear + 1 = 151 entries (Fortran near(151, n), component arrays
// alike); labels are 0-based molecule indices.
#include <string.h>

#include <vector>

#include "g6_internal.h"
#include "g6_q6.h"

#pragma clang fp contract(off)

namespace g6 {

constexpr int kMaxNear = 150;         // MAXNEAR (Neighbor.F90:4)
constexpr int kRow = kMaxNear + 1;    // entries of one array
constexpr int kNbBin = 1024;          // candidates of one molecule sorted in LDS
constexpr int kOrderNone = 0, kOrderDist = 1, kOrderLabel = 2;

// d = coord_i - coord_j - o of pair k (Neighbor.F90:119-126)
__device__ inline double nb_pair(int64_t k, const int *pi, const int *pj, const double *ox,
                                 const double *oy, const double *oz, const double *x,
                                 const double *y, const double *z, double &dx, double &dy,
                                 double &dz) {
  const int i = pi[k], j = pj[k];
  dx = x[i] - x[j];
  dy = y[i] - y[j];
  dz = z[i] - z[j];
  dx = dx - ox[k];
  dy = dy - oy[k];
  dz = dz - oz[k];
  return dx * dx + dy * dy + dz * dz;
}

// r2 < 0: no radius; otherwise radius**2 < dd drops the entry
__device__ inline bool nb_keep(double r2, double dd) { return !(r2 >= 0.0 && r2 < dd); }

__global__ void k_nb_count(int64_t np, const int *pi, const int *pj, const double *ox,
                           const double *oy, const double *oz, const double *x, const double *y,
                           const double *z, double r2, int *cnt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  double dx, dy, dz;
  if (!nb_keep(r2, nb_pair(k, pi, pj, ox, oy, oz, x, y, z, dx, dy, dz))) return;
  atomicAdd(&cnt[pi[k]], 1);
  atomicAdd(&cnt[pj[k]], 1);
}

// bin entries: 2k for the i side of pair k, 2k + 1 for the j side
__global__ void k_nb_scatter(int64_t np, const int *pi, const int *pj, const double *ox,
                             const double *oy, const double *oz, const double *x,
                             const double *y, const double *z, double r2, const int *off,
                             int *fill, long long *ent) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  double dx, dy, dz;
  if (!nb_keep(r2, nb_pair(k, pi, pj, ox, oy, oz, x, y, z, dx, dy, dz))) return;
  const int i = pi[k], j = pj[k];
  ent[off[i] + atomicAdd(&fill[i], 1)] = 2 * k;
  ent[off[j] + atomicAdd(&fill[j], 1)] = 2 * k + 1;
}

// rank of entry t among the kept entries 0..cnt-1 under (key, position)
__device__ inline int nb_rank(int cnt, const double *key, const int *seq, const int *kept,
                              int t) {
  const double kt = key[t];
  const int st = seq[t];
  int r = 0;
  for (int u = 0; u < cnt; u++) {
    const double ku = key[u];
    r += (kept[u] && (ku < kt || (ku == kt && seq[u] < st))) ? 1 : 0;
  }
  return r;
}

// neighbors_all2: one wave per molecule m sorts its bin; keys: none
// (ORDER_NONE), the distance or the label; ties by pair index
__global__ __launch_bounds__(64) void k_nb_all2(
    int n, const int *pi, const int *pj, const double *ox, const double *oy, const double *oz,
    const double *x, const double *y, const double *z, const int *off, const long long *ent,
    int maxnei, int order, int *nnear, int *near, double *ndx, double *ndy, double *ndz,
    double *ndd, int *err) {
  __shared__ double s_key[kNbBin];
  __shared__ int s_seq[kNbBin];
  __shared__ int s_one[kNbBin];
  const int m = blockIdx.x, lane = threadIdx.x;
  const int b0 = off[m], cnt = off[m + 1] - b0;
  if (cnt > kNbBin) {
    if (lane == 0) {
      atomicOr(err, 1);
      nnear[m] = 0;
    }
    return;
  }
  for (int t = lane; t < cnt; t += 64) {
    const long long e = ent[b0 + t];
    const int64_t k = e >> 1;
    double dx, dy, dz;
    const double dd = nb_pair(k, pi, pj, ox, oy, oz, x, y, z, dx, dy, dz);
    s_seq[t] = (int)k;  // one pair lists m once
    s_one[t] = 1;
    s_key[t] = order == kOrderDist ? dd
               : order == kOrderLabel ? (double)((e & 1) ? pi[k] : pj[k])
                                      : 0.0;
  }
  __syncthreads();
  // ORDER_NONE appends without a limit (Neighbor.F90:27-45)
  const int lim = (order == kOrderNone || maxnei <= 0) ? cnt : min(cnt, maxnei);
  for (int t = lane; t < cnt; t += 64) {
    const int rank = nb_rank(cnt, s_key, s_seq, s_one, t);
    if (rank >= lim) continue;
    if (rank >= kRow) {
      atomicOr(err, 1);  // the reference would write past its arrays
      continue;
    }
    const long long e = ent[b0 + t];
    const int64_t k = e >> 1;
    double dx, dy, dz;
    const double dd = nb_pair(k, pi, pj, ox, oy, oz, x, y, z, dx, dy, dz);
    const size_t o = (size_t)m * kRow + rank;
    if (e & 1) {  // j side: i as seen from j is +d
      near[o] = pi[k];
      ndx[o] = dx;
      ndy[o] = dy;
      ndz[o] = dz;
    } else {      // i side: j as seen from i is -d
      near[o] = pj[k];
      ndx[o] = -dx;
      ndy[o] = -dy;
      ndz[o] = -dz;
    }
    ndd[o] = dd;
  }
  if (lane == 0) nnear[m] = min(lim, kRow);
}

// one array staged in LDS with its sort keys
struct NbStage {
  double key[kRow], dx[kRow], dy[kRow], dz[kRow], dd[kRow];
  int near[kRow], seq[kRow], kept[kRow];
};

__device__ inline void nb_put(int r, int o, const NbStage &s, int t, int *near, double *ndx,
                              double *ndy, double *ndz, double *ndd) {
  const size_t q = (size_t)o * kRow + r;
  near[q] = s.near[t];
  ndx[q] = s.dx[t];
  ndy[q] = s.dy[t];
  ndz[q] = s.dz[t];
  ndd[q] = s.dd[t];
}

// neighbor1_compress2 of every array: one wave per array; entries past the
// radius dropped, the rest inserted in array order under the new order/limit
__global__ __launch_bounds__(64) void k_nb_compress(
    const int *inn, const int *inear, const double *idx, const double *idy, const double *idz,
    const double *idd, int maxnei, int order, double r2, int *nnear, int *near, double *ndx,
    double *ndy, double *ndz, double *ndd) {
  __shared__ NbStage s;
  __shared__ int s_nk;
  const int m = blockIdx.x, lane = threadIdx.x;
  const int cnt = inn[m];
  if (lane == 0) s_nk = 0;
  __syncthreads();
  for (int t = lane; t < cnt; t += 64) {
    const size_t q = (size_t)m * kRow + t;
    s.near[t] = inear[q];
    s.dx[t] = idx[q];
    s.dy[t] = idy[q];
    s.dz[t] = idz[q];
    s.dd[t] = idd[q];
    s.seq[t] = t;
    s.kept[t] = nb_keep(r2, idd[q]) ? 1 : 0;
    s.key[t] = order == kOrderDist ? idd[q] : order == kOrderLabel ? (double)inear[q] : 0.0;
    if (s.kept[t]) atomicAdd(&s_nk, 1);
  }
  __syncthreads();
  const int nk = s_nk;
  const int lim = (order == kOrderNone || maxnei <= 0) ? nk : min(nk, maxnei);
  for (int t = lane; t < cnt; t += 64) {
    if (!s.kept[t]) continue;
    const int rank = nb_rank(cnt, s.key, s.seq, s.kept, t);
    if (rank < lim) nb_put(rank, m, s, t, near, ndx, ndy, ndz, ndd);
  }
  if (lane == 0) nnear[m] = lim;
}

// neighbors_displacement, first pass: for each neighbour j of the moved
// molecule (staged array 1 + j), its first entry that lists the moved
// molecule (-1: none), and the output slot of the arrays that have one
__global__ __launch_bounds__(256) void k_nb_disp_mark(int target, const int *snn,
                                                      const int *snear, int *pos, int *slot,
                                                      int *nmod) {
  __shared__ int s_p[kRow];
  const int nt = snn[0];
  for (int j = threadIdx.x; j < nt; j += blockDim.x) {
    const int r = 1 + j, c = snn[r];
    int p = -1;
    for (int t = 0; t < c; t++)
      if (snear[(size_t)r * kRow + t] == target) {
        p = t;
        break;
      }
    s_p[j] = p;
    pos[j] = p;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = 1;
    for (int j = 0; j < nt; j++) slot[j] = s_p[j] >= 0 ? m++ : -1;
    *nmod = m;
  }
}

// second pass: block 0 rebuilds the moved molecule's array (every vector
// shifted by -delta), block 1 + j the array of its neighbour j (the entry of
// the moved molecule shifted by +delta); each is re-sorted stably
// (neighbors_sort, Neighbor.F90:279-333) unless its order is ORDER_NONE
__global__ __launch_bounds__(64) void k_nb_disp(
    int order, double ddx, double ddy, double ddz, const int *snn, const int *snear,
    const double *sdx, const double *sdy, const double *sdz, const double *sdd, const int *pos,
    const int *slot, int *nnear, int *near, double *ndx, double *ndy, double *ndz, double *ndd) {
  __shared__ NbStage s;
  const int b = blockIdx.x, lane = threadIdx.x;
  int o, p;
  if (b == 0) {
    o = 0;
    p = -2;  // every entry
  } else {
    o = slot[b - 1];
    p = pos[b - 1];
    if (o < 0) return;
  }
  const int cnt = snn[b];
  for (int t = lane; t < cnt; t += 64) {
    const size_t q = (size_t)b * kRow + t;
    double dx = sdx[q], dy = sdy[q], dz = sdz[q], dd = sdd[q];
    if (p == -2) {
      dx = dx - ddx;
      dy = dy - ddy;
      dz = dz - ddz;
      dd = dx * dx + dy * dy + dz * dz;
    } else if (t == p) {
      dx = dx + ddx;
      dy = dy + ddy;
      dz = dz + ddz;
      dd = dx * dx + dy * dy + dz * dz;
    }
    s.near[t] = snear[q];
    s.dx[t] = dx;
    s.dy[t] = dy;
    s.dz[t] = dz;
    s.dd[t] = dd;
    s.seq[t] = t;
    s.kept[t] = 1;
    s.key[t] = order == kOrderDist ? dd : order == kOrderLabel ? (double)s.near[t] : 0.0;
  }
  __syncthreads();
  for (int t = lane; t < cnt; t += 64) {
    const int rank = order == kOrderNone ? t : nb_rank(cnt, s.key, s.seq, s.kept, t);
    nb_put(rank, o, s, t, near, ndx, ndy, ndz, ndd);
  }
  if (lane == 0) nnear[o] = cnt;
}

// ---- umbrella order parameters of mc2 (Q6Local.F90, Triangle.F90) -------
// qlm of every bond (i, jj) within 3 A: [i][jj][m = -6..6][re, im]
__global__ void k_q6_qlm(int n, const int *nn3, const double *dx, const double *dy,
                         const double *dz, const double *dd, double *qlm, int *err) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / kMaxBond, jj = t % kMaxBond;
  if (i >= n) return;
  const int c = nn3[i];
  if (c > kMaxBond) {
    if (jj == 0) atomicOr(err, 1);  // the reference's bond arrays end at MAXBOND
    return;
  }
  if (jj >= c) return;
  const size_t e = (size_t)i * kRow + jj;
  const double pp = 1.0 / sqrt(dd[e]);
  const double px = dx[e] * pp, py = dy[e] * pp, pz = dz[e] * pp;
  double q[14];
  q6_ylm6(px, py, pz, q);
  double *o = qlm + ((size_t)i * kMaxBond + jj) * 26;
  for (int m = -6; m <= 6; m++) {
    const int am = m < 0 ? -m : m;
    double re = q[2 * am], im = q[2 * am + 1];
    if (m < 0) {
      const double sg = (am & 1) ? -1.0 : 1.0;
      re = sg * re;
      im = sg * -im;
    }
    o[2 * (m + 6)] = re;
    o[2 * (m + 6) + 1] = im;
  }
}

// label in a sorted array (binary search)
__device__ inline bool nb_has(const int *a, int c, int key) {
  int lo = 0, hi = c;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo < c && a[lo] == key;
}

// q6 of molecule k (one wave): the sum of qlm over the bonds (i, j) with i
// and j both in neis(k) = nei5(k) + k, in the order of i then of the bond,
// which is the order the reference adds them to qlmsum(k); lane c < 26 holds
// one component.  q6 = sqrt(sum_m |qlmsum|^2 * 4 pi / 13) / nbond
// (Q6Local.F90:101-231).
__global__ __launch_bounds__(64) void k_q6_sum(int n, const int *nn3, const int *near3,
                                                const int *nn5, const int *near5,
                                                const double *qlm, double *q6, int *nbond) {
  __shared__ double s_a[13];
  const int k = blockIdx.x, lane = threadIdx.x;
  const int c5 = nn5[k];
  const int *a5 = near5 + (size_t)k * kRow;
  double acc = 0.0;
  int nb = 0;
  // neis(k) in label order: nei5(k) with k inserted
  bool kdone = false;
  for (int p = 0; p <= c5; p++) {
    int i;
    if (p < c5 && (kdone || a5[p] < k)) {
      i = a5[p];
    } else if (!kdone) {
      i = k;
      kdone = true;
      p--;  // a5[p] still to come
    } else {
      break;
    }
    const int c3 = nn3[i];
    for (int jj = 0; jj < c3; jj++) {
      const int j = near3[(size_t)i * kRow + jj];
      if (j != k && !nb_has(a5, c5, j)) continue;
      nb++;
      if (lane < 26) acc = acc + qlm[((size_t)i * kMaxBond + jj) * 26 + lane];
    }
  }
  // |qlmsum(m)| for each m, then the sum over m in order
  const double re = __shfl(acc, (2 * lane) & 63, 64), im = __shfl(acc, (2 * lane + 1) & 63, 64);
  if (lane < 13) {
    const double a = hypot(re, im);
    s_a[lane] = a * a;
  }
  __syncthreads();
  if (lane == 0) {
    double nq = 0.0;
    for (int m = 0; m < 13; m++) nq = nq + s_a[m];
    q6[k] = sqrt(nq * 4.0 * kPiT / (2.0 * 6 + 1.0)) / (double)nb;
    nbond[k] = nb;
  }
}

// the sum of q6 in molecule order and op = sum / n
__global__ void k_q6_total(int n, const double *q6, double *out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; i++) s = s + q6[i];
  out[0] = s;
  out[1] = s / n;
}

// the O-O-O angles of each molecule's (up to) four nearest neighbours
// within 5 A, binned as histogram_accumulate: 21 bins of 0.1 from -1
// (Triangle.F90:53-116, DistribModule.F90:149-158); counts are integers
__global__ void k_tri_hist(int n, const int *nn, const double *dx, const double *dy,
                           const double *dz, const double *dd, unsigned int *cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t e = (size_t)i * kRow;
  const int c = nn[i];
  for (int j = 0; j < c; j++)
    for (int k = j + 1; k < c; k++) {
      double cs = dx[e + j] * dx[e + k] + dy[e + j] * dy[e + k] + dz[e + j] * dz[e + k];
      cs = cs / sqrt(dd[e + j] * dd[e + k]);
      const int bin = (int)round((cs - (-1.0)) / 0.1) + 1;
      if (1 <= bin && bin <= 21) atomicAdd(&cnt[bin - 1], 1u);
    }
}

// kullback_div of the histogram against the reference histogram
// (DistribModule.F90:304-326), and its log for the logarithmic form
__global__ void k_tri_op(const unsigned int *cnt, const double *ref, int logarithmic,
                         double *out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double h[21], s1 = 0.0, s2 = 0.0, div = 0.0;
  for (int b = 0; b < 21; b++) {
    h[b] = (double)cnt[b];
    out[b] = h[b];
  }
  for (int b = 0; b < 21; b++) {
    s1 = s1 + h[b];
    s2 = s2 + ref[b];
  }
  for (int b = 0; b < 21; b++) {
    const double p = h[b] / s1;
    double q = ref[b] / s2;
    if (q == 0) q = p;
    if (p != 0) div = div + p * log(p / q);
  }
  out[21] = logarithmic ? log(div) : div;
}

// device buffers of one call, released together
struct NbMem {
  std::vector<void *> keep;
  bool bad = false;
  template <typename T>
  T *get(size_t count) {
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(T) * (count ? count : 1)) != hipSuccess) {
      bad = true;
      return nullptr;
    }
    keep.push_back(p);
    return (T *)p;
  }
  template <typename T>
  bool put(T *d, const T *h, size_t count, hipStream_t st) {
    return count == 0 ||
           hipMemcpyAsync(d, h, sizeof(T) * count, hipMemcpyHostToDevice, st) == hipSuccess;
  }
  template <typename T>
  bool take(T *h, const T *d, size_t count, hipStream_t st) {
    return count == 0 ||
           hipMemcpyAsync(h, d, sizeof(T) * count, hipMemcpyDeviceToHost, st) == hipSuccess;
  }
  ~NbMem() {
    for (void *p : keep) (void)hipFree(p);
  }
};

// a table of nrow arrays on the device
struct NbTab {
  int *nn = nullptr, *near = nullptr;
  double *dx = nullptr, *dy = nullptr, *dz = nullptr, *dd = nullptr;
  void alloc(NbMem &mem, size_t nrow) {
    nn = mem.get<int>(nrow);
    near = mem.get<int>(nrow * kRow);
    dx = mem.get<double>(nrow * kRow);
    dy = mem.get<double>(nrow * kRow);
    dz = mem.get<double>(nrow * kRow);
    dd = mem.get<double>(nrow * kRow);
  }
  bool upload(NbMem &mem, size_t nrow, const int32_t *hnn, const int32_t *hnear,
              const double *hdx, const double *hdy, const double *hdz, const double *hdd,
              hipStream_t st) {
    const size_t e = nrow * kRow;
    return mem.put(nn, hnn, nrow, st) && mem.put(near, hnear, e, st) &&
           mem.put(dx, hdx, e, st) && mem.put(dy, hdy, e, st) && mem.put(dz, hdz, e, st) &&
           mem.put(dd, hdd, e, st);
  }
  bool download(NbMem &mem, size_t nrow, int32_t *hnn, int32_t *hnear, double *hdx,
                double *hdy, double *hdz, double *hdd, hipStream_t st) {
    const size_t e = nrow * kRow;
    return mem.take(hnn, nn, nrow, st) && mem.take(hnear, near, e, st) &&
           mem.take(hdx, dx, e, st) && mem.take(hdy, dy, e, st) && mem.take(hdz, dz, e, st) &&
           mem.take(hdd, dd, e, st);
  }
};

static bool nb_order_ok(int order) {
  return order == kOrderNone || order == kOrderDist || order == kOrderLabel;
}

// the arguments of a pair-list entry point: labels in range, no self pairs
static int nb_check_pairs(int nsite, int64_t npair, const int32_t *pi, const int32_t *pj,
                          const double *ox, const double *oy, const double *oz,
                          const double *coord) {
  if (nsite <= 0 || npair < 0 || !coord) return G6_EARG;
  if (npair > 0 && (!pi || !pj || !ox || !oy || !oz)) return G6_EARG;
  if (npair > ((int64_t)1 << 30)) return G6_EARG;
  for (int64_t k = 0; k < npair; k++)
    if (pi[k] < 0 || pi[k] >= nsite || pj[k] < 0 || pj[k] >= nsite || pi[k] == pj[k])
      return G6_EARG;
  return 0;
}

// a pair list and the molecule centres on the device
struct NbPairs {
  int n = 0;
  int64_t np = 0;
  int *pi = nullptr, *pj = nullptr;
  double *ox = nullptr, *oy = nullptr, *oz = nullptr, *x = nullptr, *y = nullptr,
         *z = nullptr;
  int upload(NbMem &mem, int nsite, int64_t npair, const int32_t *hpi, const int32_t *hpj,
             const double *hox, const double *hoy, const double *hoz, const double *coord,
             hipStream_t st) {
    n = nsite;
    np = npair;
    const size_t p = (size_t)npair, m = (size_t)nsite;
    pi = mem.get<int>(p);
    pj = mem.get<int>(p);
    ox = mem.get<double>(p);
    oy = mem.get<double>(p);
    oz = mem.get<double>(p);
    x = mem.get<double>(m);
    y = mem.get<double>(m);
    z = mem.get<double>(m);
    if (mem.bad) return G6_EHIP;
    // coord(nsite, 3): component-major, as com
    const bool ok = mem.put(x, coord, m, st) && mem.put(y, coord + m, m, st) &&
                    mem.put(z, coord + 2 * m, m, st) && mem.put(pi, hpi, p, st) &&
                    mem.put(pj, hpj, p, st) && mem.put(ox, hox, p, st) &&
                    mem.put(oy, hoy, p, st) && mem.put(oz, hoz, p, st);
    return ok ? 0 : G6_EHIP;
  }
};

// neighbors_all2 into a device table: 0, G6_ECAPACITY (an array past 151
// entries; the table is still written) or G6_EHIP
static int nb_all2_dev(NbMem &mem, hipStream_t st, const NbPairs &d, int maxnei, int order,
                       double radius, NbTab &out) {
  const size_t np = (size_t)d.np, n = (size_t)d.n;
  int *cnt = mem.get<int>(n), *off = mem.get<int>(n + 1), *fill = mem.get<int>(n);
  int *derr = mem.get<int>(1);
  long long *ent = mem.get<long long>(2 * np);
  if (mem.bad) return G6_EHIP;
  bool ok = hipMemsetAsync(cnt, 0, sizeof(int) * n, st) == hipSuccess &&
            hipMemsetAsync(fill, 0, sizeof(int) * n, st) == hipSuccess &&
            hipMemsetAsync(derr, 0, sizeof(int), st) == hipSuccess;
  if (!ok) return G6_EHIP;
  const double r2 = radius > 0 ? radius * radius : -1.0;
  const unsigned gp = (unsigned)((np + 255) / 256);
  if (np > 0)
    k_nb_count<<<gp, 256, 0, st>>>(d.np, d.pi, d.pj, d.ox, d.oy, d.oz, d.x, d.y, d.z, r2, cnt);
  std::vector<int> hc(n), ho(n + 1);
  ok = mem.take(hc.data(), cnt, n, st) && hipStreamSynchronize(st) == hipSuccess;
  if (!ok) return G6_EHIP;
  ho[0] = 0;
  for (size_t m = 0; m < n; m++) ho[m + 1] = ho[m] + hc[m];
  ok = mem.put(off, ho.data(), n + 1, st);
  if (ok && np > 0)
    k_nb_scatter<<<gp, 256, 0, st>>>(d.np, d.pi, d.pj, d.ox, d.oy, d.oz, d.x, d.y, d.z, r2, off,
                                     fill, ent);
  if (ok)
    k_nb_all2<<<d.n, 64, 0, st>>>(d.n, d.pi, d.pj, d.ox, d.oy, d.oz, d.x, d.y, d.z, off, ent,
                                  maxnei, order, out.nn, out.near, out.dx, out.dy, out.dz,
                                  out.dd, derr);
  int herr = 0;
  ok = ok && hipGetLastError() == hipSuccess && mem.take(&herr, derr, 1, st) &&
       hipStreamSynchronize(st) == hipSuccess;
  if (!ok) return G6_EHIP;
  return herr ? G6_ECAPACITY : 0;
}

// neighbor1_compress2 of every array of a device table
static void nb_compress_dev(hipStream_t st, int nrow, const NbTab &in, int maxnei, int order,
                            double radius, NbTab &out) {
  const double r2 = radius > 0 ? radius * radius : -1.0;
  k_nb_compress<<<nrow, 64, 0, st>>>(in.nn, in.near, in.dx, in.dy, in.dz, in.dd, maxnei, order,
                                     r2, out.nn, out.near, out.dx, out.dy, out.dz, out.dd);
}

}  // namespace g6

using namespace g6;

extern "C" {

int g6_neighbors(g6_handle h, int nsite, int64_t npair, const int32_t *pair_i,
                 const int32_t *pair_j, const double *ox, const double *oy, const double *oz,
                 const double *coord, int maxnei, int order, double radius, int32_t *nnear,
                 int32_t *near, double *dx, double *dy, double *dz, double *dd) {
  System *s = system_of(h);
  if (!s || !nnear || !near || !dx || !dy || !dz || !dd || maxnei < 0 || !nb_order_ok(order))
    return G6_EARG;
  int rc = nb_check_pairs(nsite, npair, pair_i, pair_j, ox, oy, oz, coord);
  if (rc) return rc;
  hipStream_t st = s->stream;
  NbMem mem;
  NbPairs d;
  NbTab out;
  out.alloc(mem, (size_t)nsite);
  if ((rc = d.upload(mem, nsite, npair, pair_i, pair_j, ox, oy, oz, coord, st))) return rc;
  const int ec = nb_all2_dev(mem, st, d, maxnei, order, radius, out);
  if (ec == G6_EHIP) return ec;
  const bool ok = out.download(mem, (size_t)nsite, nnear, near, dx, dy, dz, dd, st) &&
                  hipStreamSynchronize(st) == hipSuccess;
  if (!ok) return G6_EHIP;
  return ec;
}

int g6_neighbors_compress(g6_handle h, int nrow, const int32_t *nnear_in,
                          const int32_t *near_in, const double *dx_in, const double *dy_in,
                          const double *dz_in, const double *dd_in, int maxnei, int order,
                          double radius, int32_t *nnear, int32_t *near, double *dx, double *dy,
                          double *dz, double *dd) {
  System *s = system_of(h);
  if (!s || nrow <= 0 || !nnear_in || !near_in || !dx_in || !dy_in || !dz_in || !dd_in ||
      !nnear || !near || !dx || !dy || !dz || !dd || maxnei < 0 || !nb_order_ok(order))
    return G6_EARG;
  for (int r = 0; r < nrow; r++)
    if (nnear_in[r] < 0 || nnear_in[r] > kRow) return G6_EARG;
  const size_t n = (size_t)nrow;
  hipStream_t st = s->stream;
  NbMem mem;
  NbTab in, out;
  in.alloc(mem, n);
  out.alloc(mem, n);
  if (mem.bad) return G6_EHIP;
  bool ok = in.upload(mem, n, nnear_in, near_in, dx_in, dy_in, dz_in, dd_in, st);
  if (ok) nb_compress_dev(st, nrow, in, maxnei, order, radius, out);
  ok = ok && hipGetLastError() == hipSuccess &&
       out.download(mem, n, nnear, near, dx, dy, dz, dd, st) &&
       hipStreamSynchronize(st) == hipSuccess;
  return ok ? 0 : G6_EHIP;
}

int g6_neighbors_displacement(g6_handle h, int nsite, const int32_t *nnear_in,
                              const int32_t *near_in, const double *dx_in, const double *dy_in,
                              const double *dz_in, const double *dd_in, int order, int target,
                              double ddx, double ddy, double ddz, int32_t *nmodified,
                              int32_t *index, int32_t *nnear, int32_t *near, double *dx,
                              double *dy, double *dz, double *dd) {
  System *s = system_of(h);
  if (!s || nsite <= 0 || !nnear_in || !near_in || !dx_in || !dy_in || !dz_in || !dd_in ||
      !nmodified || !index || !nnear || !near || !dx || !dy || !dz || !dd ||
      !nb_order_ok(order) || target < 0 || target >= nsite)
    return G6_EARG;
  const int nt = nnear_in[target];
  if (nt < 0 || nt > kRow) return G6_EARG;
  // stage the arrays the move touches: the moved molecule's, then those of
  // its neighbours in its array's order
  const size_t ns = 1 + (size_t)nt;
  std::vector<int> rows(ns), snn(ns), snear(ns * kRow);
  std::vector<double> sx(ns * kRow), sy(ns * kRow), sz(ns * kRow), sd(ns * kRow);
  rows[0] = target;
  for (int j = 0; j < nt; j++) {
    const int k = near_in[(size_t)target * kRow + j];
    if (k < 0 || k >= nsite) return G6_EARG;
    rows[1 + j] = k;
  }
  for (size_t r = 0; r < ns; r++) {
    const size_t a = (size_t)rows[r] * kRow, b = r * kRow;
    snn[r] = nnear_in[rows[r]];
    if (snn[r] < 0 || snn[r] > kRow) return G6_EARG;
    memcpy(&snear[b], near_in + a, sizeof(int) * kRow);
    memcpy(&sx[b], dx_in + a, sizeof(double) * kRow);
    memcpy(&sy[b], dy_in + a, sizeof(double) * kRow);
    memcpy(&sz[b], dz_in + a, sizeof(double) * kRow);
    memcpy(&sd[b], dd_in + a, sizeof(double) * kRow);
  }
  hipStream_t st = s->stream;
  NbMem mem;
  NbTab stg, out;
  stg.alloc(mem, ns);
  out.alloc(mem, ns);
  int *pos = mem.get<int>(kRow), *slot = mem.get<int>(kRow), *dnm = mem.get<int>(1);
  if (mem.bad) return G6_EHIP;
  bool ok = stg.upload(mem, ns, snn.data(), snear.data(), sx.data(), sy.data(), sz.data(),
                       sd.data(), st);
  if (ok) {
    k_nb_disp_mark<<<1, 256, 0, st>>>(target, stg.nn, stg.near, pos, slot, dnm);
    k_nb_disp<<<(unsigned)ns, 64, 0, st>>>(order, ddx, ddy, ddz, stg.nn, stg.near, stg.dx,
                                           stg.dy, stg.dz, stg.dd, pos, slot, out.nn, out.near,
                                           out.dx, out.dy, out.dz, out.dd);
  }
  int nm = 0;
  std::vector<int> hpos(ns);
  ok = ok && hipGetLastError() == hipSuccess && mem.take(&nm, dnm, 1, st) &&
       mem.take(hpos.data(), pos, (size_t)nt, st) && hipStreamSynchronize(st) == hipSuccess;
  if (!ok || nm < 1 || nm > (int)ns) return G6_EHIP;
  ok = out.download(mem, (size_t)nm, nnear, near, dx, dy, dz, dd, st) &&
       hipStreamSynchronize(st) == hipSuccess;
  if (!ok) return G6_EHIP;
  // molecule of each output array: the moved one, then those of its
  // neighbours that list it back, in its array's order
  int m = 0;
  index[m++] = target;
  for (int j = 0; j < nt; j++)
    if (hpos[j] >= 0) index[m++] = rows[1 + j];
  *nmodified = nm;
  return m == nm ? 0 : G6_EHIP;
}

int g6_q6local(g6_handle h, int nsite, int64_t npair, const int32_t *pair_i,
               const int32_t *pair_j, const double *ox, const double *oy, const double *oz,
               const double *coord, double *q6, int32_t *nbond, double *sumq6, double *op) {
  System *s = system_of(h);
  if (!s || !q6 || !nbond || !sumq6 || !op) return G6_EARG;
  int rc = nb_check_pairs(nsite, npair, pair_i, pair_j, ox, oy, oz, coord);
  if (rc) return rc;
  hipStream_t st = s->stream;
  const size_t n = (size_t)nsite;
  NbMem mem;
  NbPairs d;
  NbTab t6, t3, t5;
  t6.alloc(mem, n);
  t3.alloc(mem, n);
  t5.alloc(mem, n);
  double *qlm = mem.get<double>(n * kMaxBond * 26), *dq6 = mem.get<double>(n),
         *tot = mem.get<double>(2);
  int *dnb = mem.get<int>(n), *derr = mem.get<int>(1);
  if (mem.bad) return G6_EHIP;
  if ((rc = d.upload(mem, nsite, npair, pair_i, pair_j, ox, oy, oz, coord, st))) return rc;
  // nei6 by label within 6 A, then nei3 and nei5 from it (Q6Local.F90:123-130)
  rc = nb_all2_dev(mem, st, d, 0, kOrderLabel, 6.0, t6);
  if (rc) return rc;
  nb_compress_dev(st, nsite, t6, 0, kOrderLabel, 3.0, t3);
  nb_compress_dev(st, nsite, t6, 0, kOrderLabel, 5.0, t5);
  if (hipMemsetAsync(derr, 0, sizeof(int), st) != hipSuccess) return G6_EHIP;
  k_q6_qlm<<<(unsigned)((n * kMaxBond + 255) / 256), 256, 0, st>>>(nsite, t3.nn, t3.dx, t3.dy,
                                                                   t3.dz, t3.dd, qlm, derr);
  int herr = 0;
  bool ok = hipGetLastError() == hipSuccess && mem.take(&herr, derr, 1, st) &&
            hipStreamSynchronize(st) == hipSuccess;
  if (!ok) return G6_EHIP;
  if (herr) return G6_ECAPACITY;  // more than MAXBOND bonds within 3 A
  k_q6_sum<<<nsite, 64, 0, st>>>(nsite, t3.nn, t3.near, t5.nn, t5.near, qlm, dq6, dnb);
  k_q6_total<<<1, 64, 0, st>>>(nsite, dq6, tot);
  double ht[2];
  ok = hipGetLastError() == hipSuccess && mem.take(q6, dq6, n, st) &&
       mem.take(nbond, dnb, n, st) && mem.take(ht, tot, 2, st) &&
       hipStreamSynchronize(st) == hipSuccess;
  if (!ok) return G6_EHIP;
  *sumq6 = ht[0];
  *op = ht[1];
  return 0;
}

int g6_triangle(g6_handle h, int nsite, int64_t npair, const int32_t *pair_i,
                const int32_t *pair_j, const double *ox, const double *oy, const double *oz,
                const double *coord, const double *refhist, int logarithmic, double *hist,
                double *op) {
  System *s = system_of(h);
  if (!s || !refhist || !hist || !op) return G6_EARG;
  int rc = nb_check_pairs(nsite, npair, pair_i, pair_j, ox, oy, oz, coord);
  if (rc) return rc;
  hipStream_t st = s->stream;
  const size_t n = (size_t)nsite;
  NbMem mem;
  NbPairs d;
  NbTab t2, t4;
  t2.alloc(mem, n);
  t4.alloc(mem, n);
  unsigned int *cnt = mem.get<unsigned int>(21);
  double *dref = mem.get<double>(21), *out = mem.get<double>(22);
  if (mem.bad) return G6_EHIP;
  if ((rc = d.upload(mem, nsite, npair, pair_i, pair_j, ox, oy, oz, coord, st))) return rc;
  // nei2 by distance within 5 A, nei = its four nearest (Triangle.F90:81-87)
  rc = nb_all2_dev(mem, st, d, 0, kOrderDist, 5.0, t2);
  if (rc) return rc;
  nb_compress_dev(st, nsite, t2, 4, kOrderDist, 0.0, t4);
  bool ok = mem.put(dref, refhist, 21, st) &&
            hipMemsetAsync(cnt, 0, sizeof(unsigned int) * 21, st) == hipSuccess;
  if (!ok) return G6_EHIP;
  k_tri_hist<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(nsite, t4.nn, t4.dx, t4.dy, t4.dz,
                                                           t4.dd, cnt);
  k_tri_op<<<1, 64, 0, st>>>(cnt, dref, logarithmic, out);
  double ho[22];
  ok = hipGetLastError() == hipSuccess && mem.take(ho, out, 22, st) &&
       hipStreamSynchronize(st) == hipSuccess;
  if (!ok) return G6_EHIP;
  for (int b = 0; b < 21; b++) hist[b] = ho[b];
  *op = ho[21];
  return 0;
}


// The order parameter after a trial displacement of molecule target by
// (dx, dy, dz), as sq6system_Differentiate / sTriangle_Differentiate return
// it: the full evaluation of the displaced centres on the same pair list.
static int displaced(int nsite, const double *coord, int target, double dx, double dy, double dz,
                     std::vector<double> &out) {
  if (!coord || target < 0 || target >= nsite || !(fabs(dx) < 1.0) || !(fabs(dy) < 1.0) ||
      !(fabs(dz) < 1.0))
    return G6_EARG;
  out.assign(coord, coord + (size_t)nsite * 3);  // coord[3][nsite]
  out[(size_t)target] += dx;
  out[(size_t)nsite + target] += dy;
  out[(size_t)2 * nsite + target] += dz;
  return 0;
}

int g6_q6local_displacement(g6_handle h, int nsite, int64_t npair, const int32_t *pair_i,
                            const int32_t *pair_j, const double *ox, const double *oy,
                            const double *oz, const double *coord, int target, double dx,
                            double dy, double dz, double *q6, int32_t *nbond, double *sumq6,
                            double *op) {
  if (nsite < 1) return G6_EARG;
  std::vector<double> c;
  const int rc = displaced(nsite, coord, target, dx, dy, dz, c);
  if (rc) return rc;
  return g6_q6local(h, nsite, npair, pair_i, pair_j, ox, oy, oz, c.data(), q6, nbond, sumq6, op);
}

int g6_triangle_displacement(g6_handle h, int nsite, int64_t npair, const int32_t *pair_i,
                             const int32_t *pair_j, const double *ox, const double *oy,
                             const double *oz, const double *coord, int target, double dx,
                             double dy, double dz, const double *refhist, int logarithmic,
                             double *hist, double *op) {
  if (nsite < 1) return G6_EARG;
  std::vector<double> c;
  const int rc = displaced(nsite, coord, target, dx, dy, dz, c);
  if (rc) return rc;
  return g6_triangle(h, nsite, npair, pair_i, pair_j, ox, oy, oz, c.data(), refhist, logarithmic,
                     hist, op);
}

}  // extern "C"
