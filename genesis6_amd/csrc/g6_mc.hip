// Monte Carlo on the GPU: the PairInfo pair-energy matrix (PairInfo.F90:12-46,
// built as mc2.F90:303-340 builds it) and the single-molecule trial loop of
// mc2 (mc2.F90:540-862, with DEBUG20050322a as the reference compiles it).
//
// The trial loop is sequential by definition, so it runs as one persistent
// block of 512 threads.  Per trial every thread derives the same proposal
// from the shared random stream (no broadcast needed), each thread tests the
// molecules it owns against the moved target (the reference's compressed
// pair list, radius rc + dt, energies by StdPairPotential), one block
// reduction forms dE in a fixed order, and every thread takes the same
// Metropolis decision.  Accepted moves update the target and its PairInfo row
// and column in place.
//
// Random numbers: the host generates the Random_GetNext stream (MT19937, the
// reference's mtrnd) and, for every position k, the Box-Muller factor
// sqrt(-2 ln u[k]) sin(2 pi u[k+1]) with the host libm, as Random_Normal
// evaluates it (RandomModule.F90:28-36).  The kernel picks the factors at the
// positions a trial consumes, so no device transcendental enters a proposal;
// only the Metropolis ratio exp() is evaluated on the device.
#include <math.h>

#include <vector>

#include "g6_internal.h"
#include "g6_mc_q6.h"
#include "g6_std.h"

#pragma clang fp contract(off)

namespace g6 {

constexpr double kI2J = 10.0, kJ2K = 0.120273;  // PhysConst.F90:48, 56
// INCREMENT=1.05 and the 0.05 floor are default-real literals in the
// reference (mc2.F90:61, 811): their single-precision values
constexpr double kIncrement = (double)1.05f, kDeltaMin = (double)0.05f;
constexpr int kMcThreads = 512;
constexpr int kMcMaxPer = 4;  // molecules per thread: n <= 2048

void make_std_table(const ModelTable &m, StdTable *t) { make_std_table(m, m, t); }

void make_std_table(const ModelTable &mi, const ModelTable &mj, StdTable *t) {
  // StdInteraction.F90:201-221: qq = q_a q_b COEFF KX2X J2I scale,
  // aa = eps_a eps_b, LJ from sqrt(aa) and the mean sigma to the 6th power
  // COEFF = CA*332.17752, KX2X = 1/X2KX, J2I = 1/I2J (PhysConst.F90:31-50);
  // site a of mi, b of mj (two models: the pairs between two components)
  const double coeff = 4.184 * 332.17752, kx2x = 1.0 / 1e-3, j2i = 1.0 / 10.0, scale = 1.0;
  t->nt = 0;
  for (int a = 0; a < mi.nsite; a++)
    for (int b = 0; b < mj.nsite; b++) {
      const double qq = mi.std_q[a] * mj.std_q[b] * coeff * kx2x * j2i * scale;
      const double aa = mi.std_eps[a] * mj.std_eps[b];
      if (qq == 0.0 && aa == 0.0) continue;
      StdTerm s{a, b, 0, qq, 0.0, 0.0};
      if (aa != 0.0) {
        const double ljeps = sqrt(aa);
        double sg = (mi.std_sig[a] + mj.std_sig[b]) * 0.5;
        sg = sg * sg * sg * sg * sg * sg;
        s.aa = 4.0 * ljeps * sg * sg * kx2x * j2i * scale;
        s.bb = -4.0 * ljeps * sg * kx2x * j2i * scale;
      }
      s.kind = qq != 0.0 ? (aa != 0.0 ? 1 : 2) : 3;
      t->t[t->nt++] = s;
    }
}

// Everything a kernel needs about the system, by value.
struct McArgs {
  int n, ns;
  double L, invL, rc, rin, c1;  // cutoff: S between rin and rc
  double beta;
  double bx[kMaxSite], by[kMaxSite], bz[kMaxSite];
  McBuf b;
  McUmb u;
};

__global__ void k_mc_sites(McArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int n = a.n;
  place_sites(a, a.b.quat[i], a.b.quat[n + i], a.b.quat[2 * n + i], a.b.quat[3 * n + i],
              a.b.com[i], a.b.com[n + i], a.b.com[2 * n + i], a.b.sx + (size_t)i * a.ns,
              a.b.sy + (size_t)i * a.ns, a.b.sz + (size_t)i * a.ns);
}

// PairInfo: one wave per molecule i computes pinfo(i,j) for j > i in the
// orientation of the all-pairs list (i < j, d = com_i - com_j), the
// compression test dd < rc^2 with the dnint shift, S and StdPairPotential,
// and writes both (i,j) and (j,i); pinfo(i,i) = 0.  part[i] = the row sum.
__global__ __launch_bounds__(64) void k_mc_pairinfo(McArgs a, StdTable st) {
  const int i = blockIdx.x, lane = threadIdx.x;
  const int n = a.n, ns = a.ns;
  const double xi = a.b.com[i], yi = a.b.com[n + i], zi = a.b.com[2 * n + i];
  const double *sxi = a.b.sx + (size_t)i * ns, *syi = a.b.sy + (size_t)i * ns,
               *szi = a.b.sz + (size_t)i * ns;
  double part = 0.0;
  for (int j = lane; j < n; j += 64) {
    double e = 0.0;
    if (j > i) {
      double d0 = xi - a.b.com[j], d1 = yi - a.b.com[n + j], d2 = zi - a.b.com[2 * n + j];
      const double k0 = round(d0 * a.invL) * a.L, k1 = round(d1 * a.invL) * a.L,
                   k2 = round(d2 * a.invL) * a.L;
      d0 = d0 - k0;
      d1 = d1 - k1;
      d2 = d2 - k2;
      const double dd = d0 * d0 + d1 * d1 + d2 * d2;
      if (dd < a.rc * a.rc) {
        const double er = mc_switch(a, dd);
        e = std_energy(st, sxi, syi, szi, a.b.sx + (size_t)j * ns, a.b.sy + (size_t)j * ns,
                       a.b.sz + (size_t)j * ns, k0, k1, k2, er);
      }
      a.b.pinfo[(size_t)i * n + j] = e;
      a.b.pinfo[(size_t)j * n + i] = e;
      part = part + e;
    } else if (j == i) {
      a.b.pinfo[(size_t)i * n + i] = 0.0;
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) part += __shfl_xor(part, d, 64);
  if (lane == 0) a.b.part[i] = part;
}

// ---- umbrellas (mc2.F90:509-521, 684-771; Triangle.F90; Umbrella.F90) ----
//
// The triangle order parameter is the Kullback-Leibler divergence of the
// histogram of O-O-O angle cosines (21 bins of 0.1 from -1), taken at every
// molecule between the vectors to its four nearest neighbours, from a
// reference histogram.  mc2 keeps, for one loop, each molecule's neighbours
// within 5 A at the loop's start (nei2) and moves only their vectors as
// molecules move (neighbors_displacement): the membership is fixed for the
// loop.  The tables here hold the same entries and vectors, unsorted; the
// nearest four are selected by distance where needed.  A trial changes the
// angles of the target and of the molecules in its table only (the vector
// to the target), so the new histogram is the old one plus, over those
// rows, the new row's angle bins minus the old row's: integer counts, the
// same numbers as the reference's set-wise update.

// keeps the four smallest keys (ascending; a key equal to a kept one goes
// behind it, as the reference's insertion in front of the first strictly
// larger key, and a full list drops it)
__device__ inline void top4_insert(double (&d)[4], int (&e)[4], double key, int idx) {
  if (!(key < d[3])) return;
  if (key < d[2]) {
    d[3] = d[2];
    e[3] = e[2];
    if (key < d[1]) {
      d[2] = d[1];
      e[2] = e[1];
      if (key < d[0]) {
        d[1] = d[0];
        e[1] = e[0];
        d[0] = key;
        e[0] = idx;
      } else {
        d[1] = key;
        e[1] = idx;
      }
    } else {
      d[2] = key;
      e[2] = idx;
    }
  } else {
    d[3] = key;
    e[3] = idx;
  }
}

// bin (0..20, or -1 outside) of the angle between two table vectors:
// costh = v_j.v_k / sqrt(dd_j dd_k), bin = nint((costh + 1) / 0.1)
// (Triangle.F90:95-103, DistribModule.F90:149-158)
__device__ inline int tri_bin(const double4 &a, const double4 &b) {
  double c = a.x * b.x + a.y * b.y + a.z * b.z;
  c = c / sqrt(a.w * b.w);
  const double r = round((c - (-1.0)) / 0.1);
  return (r >= 0.0 && r <= 20.0) ? (int)r : -1;
}

// the table entry moved by the trial: the target's own vectors lose the
// displacement, the others' vectors to the target gain it
// (neighbors_displacement, Neighbor.F90:401-457)
__device__ inline double4 tri_moved(double4 v, double sgn_dx, double sgn_dy, double sgn_dz) {
  v.x = v.x + sgn_dx;
  v.y = v.y + sgn_dy;
  v.z = v.z + sgn_dz;
  v.w = v.x * v.x + v.y * v.y + v.z * v.z;
  return v;
}

// the six (or fewer) angle bins of a row's nearest four, added to h with
// weight w (LDS integer atomics: exact, order free)
__device__ inline void tri_row_add(int *h, const double4 (&v)[4], int k, int w) {
  for (int p = 0; p < k; p++)
    for (int q = p + 1; q < k; q++) {
      const int bin = tri_bin(v[p], v[q]);
      if (bin >= 0) atomicAdd(&h[bin], w);
    }
}

// sTriangle_prepareall (Triangle.F90:53-116) for molecule i: its entries
// within 5 A (the loop's pair list within rc, neighbors_all2 radius 5: kept
// when dd <= 25), the vector v = c_j - c_i less the nearest image shift, as
// the reference forms it from the pair (i, j) in either orientation, and
// its nearest four's angle bins added to the histogram
__global__ __launch_bounds__(256) void k_mc_umb_prepare(McArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int n = a.n;
  const double xi = a.b.com[i], yi = a.b.com[n + i], zi = a.b.com[2 * n + i];
  const double rr = a.rc * a.rc;
  int cnt = 0;
  double dk[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  int ek[4] = {-1, -1, -1, -1};
  double4 *row = reinterpret_cast<double4 *>(a.b.tri_v) + (size_t)i * kTriNear;
  for (int j = 0; j < n; j++) {
    if (j == i) continue;
    double d0 = a.b.com[j] - xi, d1 = a.b.com[n + j] - yi, d2 = a.b.com[2 * n + j] - zi;
    d0 = d0 - round(d0 * a.invL) * a.L;
    d1 = d1 - round(d1 * a.invL) * a.L;
    d2 = d2 - round(d2 * a.invL) * a.L;
    const double dd = d0 * d0 + d1 * d1 + d2 * d2;
    if (!(dd < rr) || 25.0 < dd) continue;
    if (cnt < kTriNear) {
      a.b.tri_near[(size_t)i * kTriNear + cnt] = j;
      row[cnt] = make_double4(d0, d1, d2, dd);
      top4_insert(dk, ek, dd, cnt);
    }
    cnt++;
  }
  if (cnt > kTriNear) atomicCAS(a.b.uflag, 0, 1);
  const int m = cnt < kTriNear ? cnt : kTriNear;
  a.b.tri_nn[i] = m;
  const int k = m < 4 ? m : 4;
  double4 v[4];
  for (int p = 0; p < k; p++) v[p] = row[ek[p]];
  tri_row_add(a.b.tri_hist, v, k, 1);
}

// sq6system_prepareall (Q6Local.F90:101-231) for molecule i: its row of the
// molecules within 6 A (the loop's pairs within rc, neighbors_all2 radius 6,
// by label), neis(i) = those within 5 A then i, the qlm of its bonds (within
// 3 A) and its bond count.  Flag 2: a row past 64 entries or more than
// MAXBOND bonds.
__global__ __launch_bounds__(256) void k_q6u_rows(McArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int n = a.n;
  const Q6Buf &q = a.b.q6;
  const double xi = a.b.com[i], yi = a.b.com[n + i], zi = a.b.com[2 * n + i];
  const double rr = a.rc * a.rc;
  double4 *row = reinterpret_cast<double4 *>(q.v6) + (size_t)i * kQ6Near;
  int cnt = 0;
  for (int j = 0; j < n; j++) {
    if (j == i) continue;
    double d0 = a.b.com[j] - xi, d1 = a.b.com[n + j] - yi, d2 = a.b.com[2 * n + j] - zi;
    d0 = d0 - round(d0 * a.invL) * a.L;
    d1 = d1 - round(d1 * a.invL) * a.L;
    d2 = d2 - round(d2 * a.invL) * a.L;
    const double dd = d0 * d0 + d1 * d1 + d2 * d2;
    if (!(dd < rr) || 36.0 < dd) continue;
    if (cnt < kQ6Near) {
      q.lab6[(size_t)i * kQ6Near + cnt] = j;
      row[cnt] = make_double4(d0, d1, d2, dd);
    }
    cnt++;
  }
  const int m = cnt < kQ6Near ? cnt : kQ6Near;
  q.nn6[i] = m;
  int ns = 0, nb = 0;
  int *neis = q.neis + (size_t)i * kQ6Set;
  for (int e = 0; e < m; e++) {
    const double4 v = row[e];
    if (q6_is_shell(v.w)) neis[ns++] = q.lab6[(size_t)i * kQ6Near + e];
    if (q6_is_bond(v.w)) {
      if (nb < kMaxBond) q6_record(v, q6_rec(q, i, nb));
      nb++;
    }
  }
  neis[ns++] = i;
  q.neis_n[i] = ns;
  q.nb3[i] = nb;
  if (cnt > kQ6Near || nb > kMaxBond) atomicCAS(a.b.uflag, 0, 2);
}

// neib(jj, i) = neis(i) & neis(j), in neis(i)'s order, for bond jj of i
__global__ __launch_bounds__(256) void k_q6u_neib(McArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / kMaxBond, jj = t % kMaxBond;
  if (i >= a.n) return;
  const Q6Buf &q = a.b.q6;
  if (jj >= q.nb3[i] || q.nb3[i] > kMaxBond) return;
  const double4 *row = reinterpret_cast<const double4 *>(q.v6) + (size_t)i * kQ6Near;
  int j = -1, b = 0;
  for (int e = 0; e < q.nn6[i]; e++)
    if (q6_is_bond(row[e].w)) {
      if (b == jj) {
        j = q.lab6[(size_t)i * kQ6Near + e];
        break;
      }
      b++;
    }
  if (j < 0) return;
  const int *si = q.neis + (size_t)i * kQ6Set, *sj = q.neis + (size_t)j * kQ6Set;
  const int ni = q.neis_n[i], nj = q.neis_n[j];
  int *out = q6_neib(q, i, jj), c = 0;
  for (int p = 0; p < ni; p++) {
    bool in = false;
    for (int r = 0; r < nj; r++) in = in || sj[r] == si[p];
    if (in) out[c++] = si[p];
  }
  *q6_neib_n(q, i, jj) = c;
}

// qlmsum(k) and nbond(k): the bonds (i, jj) whose set holds k, added in the
// reference's order (i ascending, then jj); such an i is k or a molecule of
// k's 5 A shell.  One wave per molecule, lane c < 26 one component; q6(k).
__global__ __launch_bounds__(64) void k_q6u_sum(McArgs a) {
  const int k = blockIdx.x, lane = threadIdx.x;
  const Q6Buf &q = a.b.q6;
  const WRow rk = wr_load(q, k, lane);
  const unsigned long long sh = __ballot(lane < rk.n && q6_is_shell(rk.v.w));
  const int ns = __popcll(sh);
  double acc = 0.0;
  int nb = 0;
  bool kdone = false;
  for (int p = 0; p <= ns; p++) {
    int i;
    const int lab = p < ns ? __shfl(rk.lab, q6_select_bit(sh, p), 64) : 0x7fffffff;
    if (!kdone && k < lab) {
      i = k;
      kdone = true;
      p--;
    } else if (p < ns) {
      i = lab;
    } else {
      break;
    }
    const int nbi = q.nb3[i] < kMaxBond ? q.nb3[i] : kMaxBond;
    for (int jj = 0; jj < nbi; jj++) {
      const WSet s = ws_load(q6_neib(q, i, jj), q6_neib_n(q, i, jj), lane);
      if (!ws_has(s, k, lane)) continue;
      nb++;
      if (lane < 26) acc = acc + q6_rec(q, i, jj)[lane];
    }
  }
  if (lane < 26) q.qsum[(size_t)k * kQ6Rec + lane] = acc;
  if (lane == 0) q.nbond[k] = nb;
}

// q6 of every molecule and the order parameter sum / n
__global__ __launch_bounds__(256) void k_q6u_q6(McArgs a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < a.n) a.b.q6.dq6[k] = q6_of(a.b.q6, k);
}
__global__ void k_q6u_total(McArgs a) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double s = 0.0;
  for (int k = 0; k < a.n; k++) s = s + a.b.q6.dq6[k];
  *a.b.q6.op = s / a.n;
}

// kullback_div (DistribModule.F90:304-326) of h (counts) from ref
__device__ inline double kl_div(const int *h, const double *ref) {
  double s1 = 0.0, s2 = 0.0, div = 0.0;
  for (int b = 0; b < 21; b++) {
    s1 = s1 + (double)h[b];
    s2 = s2 + ref[b];
  }
  for (int b = 0; b < 21; b++) {
    const double p = (double)h[b] / s1;
    double q = ref[b] / s2;
    if (q == 0) q = p;
    if (p != 0) div = div + p * log(p / q);
  }
  return div;
}

__device__ inline double umbrella_potential(double coeffi, double center, double op) {
  return coeffi * ((op - center) * (op - center));
}

// mc2's trial loop, ntrial trials, one block.  The configuration (COM,
// quaternion, step sizes) lives in LDS for the whole call, and the random
// stream is staged into LDS kMcChunk trials at a time; the only global
// traffic per trial is the PairInfo row of the target (issued first) and,
// on acceptance, the PairInfo and configuration write-back.
constexpr int kMcChunk = 192;                 // trials per staged chunk
constexpr int kMcStage = kMcChunk * 14 + 2;   // stream entries per chunk
template <int M, bool UQ6>
__global__ __launch_bounds__(kMcThreads) void k_mc_trials(McArgs a, StdTable st, int ntrial) {
  constexpr int NS = StdPattern<M>::NS;
  extern __shared__ __align__(16) double s_mem[];
  const int n = a.n;
  double *l_com = s_mem;                      // [3][n]
  double *l_q = l_com + 3 * n;                // [4][n]
  double *l_delta = l_q + 4 * n;              // [n]
  double *l_rtr = l_delta + n;                // [n]
  double *l_u = l_rtr + n;                    // [kMcStage]
  double *l_z = l_u + kMcStage;               // [kMcStage]
  __shared__ double s_red[kMcThreads / 64];
  // umbrella state: histogram, the trial's change, reference, order
  // parameters now and after the trial, the trial's total energy change
  __shared__ int s_hist[21], s_dh[21], s_abort;
  __shared__ double s_ref[21], s_op[kUmbMax], s_newop[kUmbMax], s_des;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int numb = a.u.numb;
  if (numb > 0 && *a.b.uflag != 0) return;  // the tables did not fit
  for (int k = tid; k < 3 * n; k += kMcThreads) l_com[k] = a.b.com[k];
  for (int k = tid; k < 4 * n; k += kMcThreads) l_q[k] = a.b.quat[k];
  for (int k = tid; k < n; k += kMcThreads) {
    l_delta[k] = a.b.delta[k];
    l_rtr[k] = a.b.rtr[k];
  }
  int pos = 0;    // position in the call's stream
  int base = -1;  // stream position of l_u[0]
  long long nacc = 0;
  double epsum = *a.b.epsum;
  if (numb > 0) {
    if (tid < 21) {
      s_hist[tid] = a.u.tri ? a.b.tri_hist[tid] : 0;
      s_dh[tid] = 0;
      s_ref[tid] = a.u.refhist[tid];
    }
    if (tid == 0) s_abort = 0;
    __syncthreads();
    if (tid == 0) {  // prepareall's op (mc2.F90:509-521)
      const double div = a.u.tri ? kl_div(s_hist, s_ref) : 0.0;
      for (int k = 0; k < numb; k++) {
        const int kind = a.u.kind[k];
        s_op[k] = kind == 3 ? epsum : kind == 4 ? log(div) : kind == 2 ? *a.b.q6.op : div;
      }
    }
  }
  bool aborted = false;
  for (int trial = 0; trial < ntrial; trial++) {
    if (trial % kMcChunk == 0) {  // stage the next chunk (a trial reads <= 14)
      __syncthreads();
      base = pos;
      for (int k = tid; k < kMcStage; k += kMcThreads) {
        l_u[k] = a.b.u[base + k];
        l_z[k] = a.b.z[base + k];
      }
      __syncthreads();
    }
    const int p = pos - base;
    const int trial_pos = pos;
    // target = random_getnext*nmol + 1 (1-based); u = 1 exactly would name
    // molecule n+1, which the reference does not guard: clamped here
    int target = (int)(l_u[p] * (double)n + 1.0) - 1;
    if (target >= n) target = n - 1;
    double pold[kMcMaxPer];
#pragma unroll
    for (int m = 0; m < kMcMaxPer; m++) {
      const int j = tid + m * kMcThreads;
      pold[m] = j < n ? a.b.pinfo[(size_t)target * n + j] : 0.0;
    }
    const double dt = l_delta[target];
    const double dr = l_delta[target] * l_rtr[target];
    const double rotx = l_z[p + 1] * sqrt(dr * dr) + 0.0;
    const double roty = l_z[p + 3] * sqrt(dr * dr) + 0.0;
    const double rotz = l_z[p + 5] * sqrt(dr * dr) + 0.0;
    const double avgrot = (fabs(rotx) + fabs(roty) + fabs(rotz)) / 3.0;
    // qinfrotator (QuatModule.F90:203-211) + qnormalize (215-223)
    double qa = 1.0, qb = rotx, qc = roty, qd = rotz;
    {
      const double r = 1.0 / sqrt(qa * qa + qb * qb + qc * qc + qd * qd);
      qa = qa * r;
      qb = qb * r;
      qc = qc * r;
      qd = qd * r;
    }
    // qadd (QuatModule.F90:108-133): new = rotator x old, a >= 0
    {
      const double a2 = l_q[target], b2 = l_q[n + target], c2 = l_q[2 * n + target],
                   d2 = l_q[3 * n + target];
      const double a3 = qa * a2 - qb * b2 - qc * c2 - qd * d2;
      const double b3 = qa * b2 + qb * a2 + qc * d2 - qd * c2;
      const double c3 = qa * c2 + qc * a2 - qb * d2 + qd * b2;
      const double d3 = qa * d2 + qd * a2 + qb * c2 - qc * b2;
      qa = a3;
      qb = b3;
      qc = c3;
      qd = d3;
      if (qa < 0) {
        qa = -qa;
        qb = -qb;
        qc = -qc;
        qd = -qd;
      }
    }
    const double deltax = l_z[p + 7] * sqrt(dt * dt) + 0.0;
    const double deltay = l_z[p + 9] * sqrt(dt * dt) + 0.0;
    const double deltaz = l_z[p + 11] * sqrt(dt * dt) + 0.0;
    const double avgtrans = (fabs(deltax) + fabs(deltay) + fabs(deltaz)) / 3.0;
    pos += 13;
    const double tx = deltax + l_com[target], ty = deltay + l_com[n + target],
                 tz = deltaz + l_com[2 * n + target];
    double sxt[kMaxSite], syt[kMaxSite], szt[kMaxSite];
    place_sites_t<NS>(a, qa, qb, qc, qd, tx, ty, tz, sxt, syt, szt);
    // the compressed list of the moved target (radius rc + dt) and dE; the
    // sites of a listed molecule are placed again from its COM and quaternion
    const double rr = (a.rc + dt) * (a.rc + dt);
    double enew[kMcMaxPer], local = 0.0;
#pragma unroll
    for (int m = 0; m < kMcMaxPer; m++) {
      const int j = tid + m * kMcThreads;
      enew[m] = 0.0;
      if (j < n) {
        if (j != target) {
          double d0 = l_com[j] - tx, d1 = l_com[n + j] - ty, d2 = l_com[2 * n + j] - tz;
          const double k0 = round(d0 * a.invL) * a.L, k1 = round(d1 * a.invL) * a.L,
                       k2 = round(d2 * a.invL) * a.L;
          d0 = d0 - k0;
          d1 = d1 - k1;
          d2 = d2 - k2;
          const double dd = d0 * d0 + d1 * d1 + d2 * d2;
          if (dd < rr) {
            const double er = mc_switch(a, dd);
            double xj[kMaxSite], yj[kMaxSite], zj[kMaxSite];
            place_sites_t<NS>(a, l_q[j], l_q[n + j], l_q[2 * n + j], l_q[3 * n + j], l_com[j],
                              l_com[n + j], l_com[2 * n + j], xj, yj, zj);
            enew[m] = std_energy_t<M>(st, xj, yj, zj, sxt, syt, szt, k0, k1, k2, er);
          }
        }
        local = local + (enew[m] - pold[m]);
      }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) local += __shfl_xor(local, d, 64);
    if (lane == 0) s_red[wid] = local;
    // the triangle rows the trial changes: the target (tid 0) and the
    // molecules of its table (tid 1..nt); tslot = the row's entry of the
    // target, whose vector the move changes
    int trow = -1, tslot = -1;
    if (a.u.tri) {
      const int nt = a.b.tri_nn[target];
      if (tid <= nt) {
        trow = tid == 0 ? target : a.b.tri_near[(size_t)target * kTriNear + tid - 1];
        const int m = a.b.tri_nn[trow];
        const double4 *row = reinterpret_cast<const double4 *>(a.b.tri_v) + (size_t)trow * kTriNear;
        const int *lab = a.b.tri_near + (size_t)trow * kTriNear;
        double dold[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
        double dnew[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
        int eold[4] = {-1, -1, -1, -1}, enew4[4] = {-1, -1, -1, -1};
        for (int e = 0; e < m; e++) {
          const double4 v = row[e];
          top4_insert(dold, eold, v.w, e);
          double nd = v.w;
          if (tid == 0) {
            nd = tri_moved(v, -deltax, -deltay, -deltaz).w;
          } else if (lab[e] == target) {
            nd = tri_moved(v, deltax, deltay, deltaz).w;
            tslot = e;
          }
          top4_insert(dnew, enew4, nd, e);
        }
        const int k = m < 4 ? m : 4;
        double4 v[4];
        for (int p = 0; p < k; p++) v[p] = row[eold[p]];
        tri_row_add(s_dh, v, k, -1);
        for (int p = 0; p < k; p++) {
          v[p] = row[enew4[p]];
          if (tid == 0)
            v[p] = tri_moved(v[p], -deltax, -deltay, -deltaz);
          else if (enew4[p] == tslot)
            v[p] = tri_moved(v[p], deltax, deltay, deltaz);
        }
        tri_row_add(s_dh, v, k, 1);
      }
    }
    // the Q6 umbrella's Differentiate on wave 0: the target's row before
    // and after the move; a move that would give a molecule more than
    // MAXBOND bonds stops the call before it changes anything
    [[maybe_unused]] Q6Trial qt;
    if constexpr (UQ6) {
      if (wid == 0) {
        const Q6Buf &q = a.b.q6;
        qt.old_t = wr_load(q, target, lane);
        qt.old_b = wr_bonds(qt.old_t, lane);
        qt.new_t = qt.old_t;
        if (lane < qt.new_t.n) qt.new_t.v = tri_moved(qt.old_t.v, -deltax, -deltay, -deltaz);
        qt.new_b = wr_bonds(qt.new_t, lane);
        if (!q6_bonds_fit(q, qt, lane)) {
          if (lane == 0) s_abort = 1;
        } else {
          q6_differentiate(q, qt, target, lane);
        }
      }
    }
    __syncthreads();
    if constexpr (UQ6) {
      if (s_abort) {  // uniform: every thread leaves before the decision
        pos = trial_pos;
        aborted = true;
        break;
      }
      for (int k = tid; k < n; k += kMcThreads) a.b.q6.dq6[k] = q6_of(a.b.q6, k);
      __syncthreads();
    }
    double de = 0.0;
#pragma unroll
    for (int w = 0; w < kMcThreads / 64; w++) de = de + s_red[w];
    const double deltae = de;
    if (numb > 0) {
      // deltaesum = deltae + sum(U(newop) - U(op)) (mc2.F90:684-706),
      // formed once and shared
      if (tid == 0) {
        double div = 0.0;
        if (a.u.tri) {
          int h[21];
          for (int b = 0; b < 21; b++) h[b] = s_hist[b] + s_dh[b];
          div = kl_div(h, s_ref);
        }
        double q6op = 0.0;
        if constexpr (UQ6) {  // sumq6 / nmol (Q6Local.F90:429-440)
          double sq = 0.0;
          for (int k = 0; k < n; k++) sq = sq + a.b.q6.dq6[k];
          q6op = sq / n;
        }
        double dope = 0.0;
        for (int k = 0; k < numb; k++) {
          const int kind = a.u.kind[k];
          const double nop = kind == 3 ? epsum + deltae
                             : kind == 4 ? log(div)
                             : kind == 2 ? q6op
                                         : div;
          s_newop[k] = nop;
          dope = dope + umbrella_potential(a.u.coeffi[k], a.u.center[k], nop) -
                 umbrella_potential(a.u.coeffi[k], a.u.center[k], s_op[k]);
        }
        s_des = deltae + dope;
      }
      __syncthreads();
      de = s_des;
    }
    bool accept;
    if (de < 0.0) {
      accept = true;
    } else {
      const double ratio = exp(-(de * kI2J * kJ2K) * a.beta);
      accept = l_u[pos - base] < ratio;
      pos += 1;
    }
    if (numb > 0) {
      if (accept) {
        if (tid < numb) s_op[tid] = s_newop[tid];
        if (tid < 21) s_hist[tid] = s_hist[tid] + s_dh[tid];
        // sTriangle_Integrate: the moved vectors into the tables
        if (trow >= 0) {
          double4 *row = reinterpret_cast<double4 *>(a.b.tri_v);
          if (tid > 0) {
            const size_t et = (size_t)target * kTriNear + tid - 1;
            row[et] = tri_moved(row[et], -deltax, -deltay, -deltaz);
            if (tslot >= 0) {
              const size_t es = (size_t)trow * kTriNear + tslot;
              row[es] = tri_moved(row[es], deltax, deltay, deltaz);
            }
          }
        }
      }
      if (tid < 21) s_dh[tid] = 0;
    }
    if constexpr (UQ6) {  // sq6system_Integrate or _Reject
      if (wid == 0) {
        int ovf = 0;
        if (accept)
          q6_integrate(a.b.q6, qt, target, deltax, deltay, deltaz, lane, ovf);
        else
          q6_reject(a.b.q6, qt, target, lane);
      }
    }
    if (accept) {
      nacc += 1;
      epsum = epsum + deltae;
#pragma unroll
      for (int m = 0; m < kMcMaxPer; m++) {
        const int j = tid + m * kMcThreads;
        if (j < n) {
          a.b.pinfo[(size_t)j * n + target] = enew[m];
          a.b.pinfo[(size_t)target * n + j] = enew[m];
        }
      }
      if (tid == 0) {
        l_com[target] = tx;
        l_com[n + target] = ty;
        l_com[2 * n + target] = tz;
        l_q[target] = qa;
        l_q[n + target] = qb;
        l_q[2 * n + target] = qc;
        l_q[3 * n + target] = qd;
        l_delta[target] = l_delta[target] * kIncrement;
        if (avgtrans * l_rtr[target] < avgrot)
          l_rtr[target] = l_rtr[target] * kIncrement;
        else
          l_rtr[target] = l_rtr[target] / kIncrement;
      } else if (tid == 64) {  // the site positions the host may read back
#pragma unroll
        for (int s = 0; s < NS; s++) {
          a.b.sx[(size_t)target * a.ns + s] = sxt[s];
          a.b.sy[(size_t)target * a.ns + s] = syt[s];
          a.b.sz[(size_t)target * a.ns + s] = szt[s];
        }
      }
    } else if (tid == 0) {
      l_delta[target] = l_delta[target] / kIncrement;
      if (l_delta[target] < kDeltaMin) l_delta[target] = kDeltaMin;
    }
    __syncthreads();
  }
  for (int k = tid; k < 3 * n; k += kMcThreads) a.b.com[k] = l_com[k];
  for (int k = tid; k < 4 * n; k += kMcThreads) a.b.quat[k] = l_q[k];
  for (int k = tid; k < n; k += kMcThreads) {
    a.b.delta[k] = l_delta[k];
    a.b.rtr[k] = l_rtr[k];
  }
  if (tid == 0) {
    *a.b.pos = pos;
    *a.b.nacc = *a.b.nacc + nacc;
    *a.b.epsum = epsum;
    for (int k = 0; k < numb; k++) a.b.uop[k] = s_op[k];
    if (aborted) *a.b.uflag = 3;
  }
}

// LDS bytes of k_mc_trials for n molecules
static size_t mc_lds(int n) { return sizeof(double) * ((size_t)9 * n + 2 * kMcStage); }

static McArgs mc_args(const System &s) {
  McArgs a{};
  a.n = s.n;
  a.ns = s.ns;
  a.L = s.L;
  a.invL = 1.0 / s.L;
  a.rc = s.rc;
  a.rin = s.rc - s.taper;
  a.c1 = s.taper > 0 ? s.geo.c1 : 0.0;
  a.beta = s.mc_beta;
  for (int k = 0; k < kMaxSite; k++) {
    a.bx[k] = s.mt.bx[k];
    a.by[k] = s.mt.by[k];
    a.bz[k] = s.mt.bz[k];
  }
  a.b = s.mc;
  a.u = s.mc_umb;
  return a;
}

void mc_free(System &s) {
  McBuf &m = s.mc;
  Q6Buf &q = m.q6;
  void *ps[] = {m.com, m.quat, m.sx, m.sy, m.sz, m.pinfo, m.delta, m.rtr, m.u, m.z,
                m.epsum, m.part, m.pos, m.nacc, m.tri_nn, m.tri_near, m.tri_hist,
                m.uflag, m.tri_v, m.uop, q.nn6, q.lab6, q.v6, q.neis_n, q.neis, q.neib_n,
                q.neib, q.qlm, q.qsum, q.dq6, q.op, q.nbond, q.nb3};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  m = McBuf{};
  s.mc_umb = McUmb{};
  s.mc_ready = false;
}

void McHost::seed(uint32_t s) {
  mt[0] = s;
  for (int i = 1; i < 624; i++) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
  mti = 624;
}

uint32_t McHost::next() {
  if (mti >= 624) {
    for (int k = 0; k < 624; k++) {
      const uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
      mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    mti = 0;
  }
  uint32_t y = mt[mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

double McHost::getnext() { return 1.0 - (double)(next() >> 1) * (1.0 / 2147483648.0); }

}  // namespace g6

using namespace g6;

extern "C" {

int g6_mc_begin(g6_handle h, double temp, uint32_t seed, const double *delta,
                const double *rtratio) {
  System *s = system_of(h);
  if (!s || !(temp > 0) || !delta || !rtratio) return G6_EARG;
  if (!is_cubic(*s)) return G6_EARG;  // mc2's trial loop runs cubic boxes here
  const int n = s->n, ns = s->ns;
  if (n > kMcThreads * kMcMaxPer || mc_lds(n) > 150 * 1024) return G6_EARG;
  mc_free(*s);
  McBuf &m = s->mc;
  auto A = [](void **p, size_t b) { return hipMalloc(p, b > 0 ? b : 8); };
#define AL(ptr, bytes) G6_CHECK(A((void **)&(ptr), (bytes)))
  AL(m.com, sizeof(double) * 3 * n);
  AL(m.quat, sizeof(double) * 4 * n);
  AL(m.sx, sizeof(double) * (size_t)n * ns);
  AL(m.sy, sizeof(double) * (size_t)n * ns);
  AL(m.sz, sizeof(double) * (size_t)n * ns);
  AL(m.pinfo, sizeof(double) * (size_t)n * n);
  AL(m.delta, sizeof(double) * n);
  AL(m.rtr, sizeof(double) * n);
  AL(m.epsum, sizeof(double));
  AL(m.part, sizeof(double) * n);
  AL(m.pos, sizeof(int));
  AL(m.nacc, sizeof(long long));
#undef AL
  // the configuration of the MD state: COM and the order-1 quaternion
  G6_CHECK(hipMemcpy(m.com, s->md.com, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice));
  G6_CHECK(hipMemcpy(m.quat, s->md.quat, sizeof(double) * 4 * n, hipMemcpyDeviceToDevice));
  G6_CHECK(hipMemcpy(m.delta, delta, sizeof(double) * n, hipMemcpyHostToDevice));
  G6_CHECK(hipMemcpy(m.rtr, rtratio, sizeof(double) * n, hipMemcpyHostToDevice));
  G6_CHECK(hipMemset(m.nacc, 0, sizeof(long long)));
  make_std_table(s->mt, &s->std);
  bool match = false;
  std_model(s->model, [&](auto m) { match = std_table_matches<decltype(m)::value>(s->std); });
  if (!match) return G6_EARG;
  s->mc_beta = 1.0 / temp;  // @TEMP: beta = 1/temp (mc2util.F90)
  McHost *hst = &s->mch;
  hst->seed(seed);
  hst->pend.clear();
  s->mc_ready = true;
  McArgs a = mc_args(*s);
  k_mc_sites<<<(n + 255) / 256, 256, 0, s->stream>>>(a);
  k_mc_pairinfo<<<n, 64, 0, s->stream>>>(a, s->std);
  G6_CHECK(hipGetLastError());
  std::vector<double> part(n);
  G6_CHECK(hipMemcpyAsync(part.data(), m.part, sizeof(double) * n, hipMemcpyDeviceToHost,
                          s->stream));
  G6_CHECK(hipStreamSynchronize(s->stream));
  double e = 0.0;
  for (int i = 0; i < n; i++) e = e + part[i];
  G6_CHECK(hipMemcpy(m.epsum, &e, sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int g6_mc_trials(g6_handle h, int ntrial, long long *naccept) {
  System *s = system_of(h);
  if (!s || !s->mc_ready || ntrial < 0) return G6_EARG;
  if (ntrial == 0) return 0;
  McBuf &m = s->mc;
  McHost *hst = &s->mch;
  // a trial consumes 13 or 14 numbers; the factor at k also reads k + 1;
  // the kernel stages whole chunks, so the last one may read kMcStage past
  // the position it starts at
  const size_t need = (size_t)ntrial * 14 + 2 + kMcStage;
  while (hst->pend.size() < need) hst->pend.push_back(hst->getnext());
  if ((int)need > m.ucap) {
    if (m.u) (void)hipFree(m.u);
    if (m.z) (void)hipFree(m.z);
    m.u = m.z = nullptr;
    G6_CHECK(hipMalloc(&m.u, sizeof(double) * need));
    G6_CHECK(hipMalloc(&m.z, sizeof(double) * need));
    m.ucap = (int)need;
  }
  std::vector<double> z(need, 0.0);
  const double pi = 3.14159265358979323846;  // PhysConst.F90 PI
  for (size_t k = 0; k + 1 < need; k++) {
    const double r = sqrt(-2.0 * log(hst->pend[k]));
    const double theta = 2.0 * pi * hst->pend[k + 1];
    z[k] = r * sin(theta);
  }
  G6_CHECK(hipMemcpy(m.u, hst->pend.data(), sizeof(double) * need, hipMemcpyHostToDevice));
  G6_CHECK(hipMemcpy(m.z, z.data(), sizeof(double) * need, hipMemcpyHostToDevice));
  McArgs a = mc_args(*s);
  if (s->mc_umb.numb > 0) {
    // the order parameters of the loop's start (mc2.F90:486-521)
    G6_CHECK(hipMemsetAsync(m.uflag, 0, sizeof(int), s->stream));
    if (s->mc_umb.tri) {
      G6_CHECK(hipMemsetAsync(m.tri_hist, 0, sizeof(int) * 21, s->stream));
      k_mc_umb_prepare<<<(s->n + 255) / 256, 256, 0, s->stream>>>(a);
      G6_CHECK(hipGetLastError());
    }
    if (s->mc_umb.q6) {
      const int n = s->n;
      k_q6u_rows<<<(n + 255) / 256, 256, 0, s->stream>>>(a);
      k_q6u_neib<<<(n * kMaxBond + 255) / 256, 256, 0, s->stream>>>(a);
      k_q6u_sum<<<n, 64, 0, s->stream>>>(a);
      k_q6u_q6<<<(n + 255) / 256, 256, 0, s->stream>>>(a);
      k_q6u_total<<<1, 1, 0, s->stream>>>(a);
      G6_CHECK(hipGetLastError());
    }
  }
  const size_t lds = mc_lds(s->n);
  hipError_t attr = hipSuccess;
  auto launch = [&](auto kern) {
    attr = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds);
    if (attr == hipSuccess) kern<<<1, kMcThreads, lds, s->stream>>>(a, s->std, ntrial);
  };
  std_model(s->model, [&](auto m) {
    constexpr int M = decltype(m)::value;
    if (s->mc_umb.q6)
      launch(k_mc_trials<M, true>);
    else
      launch(k_mc_trials<M, false>);
  });
  G6_CHECK(attr);
  G6_CHECK(hipGetLastError());
  int used = 0, uflag = 0;
  long long nacc = 0;
  if (s->mc_umb.numb > 0)
    G6_CHECK(hipMemcpyAsync(&uflag, m.uflag, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  G6_CHECK(hipMemcpyAsync(&used, m.pos, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  G6_CHECK(hipMemcpyAsync(&nacc, m.nacc, sizeof(long long), hipMemcpyDeviceToHost, s->stream));
  G6_CHECK(hipStreamSynchronize(s->stream));
  // flags 1, 2: the loop's tables did not fit and nothing ran (the stream is
  // kept whole); 3: a move would have given a molecule more than MAXBOND
  // bonds, and the trials before it stand
  if (uflag == 1 || uflag == 2) return G6_ECAPACITY;
  if (used < 0 || (size_t)used > need) return G6_EHIP;
  hst->pend.erase(hst->pend.begin(), hst->pend.begin() + used);
  if (naccept) *naccept = nacc;
  return uflag ? G6_ECAPACITY : 0;
}

int g6_mc_set_umbrella(g6_handle h, int numbrella, const int32_t *kind, const double *center,
                       const double *coeffi, const double *refhist) {
  System *s = system_of(h);
  if (!s || !s->mc_ready || numbrella < 0 || numbrella > kUmbMax) return G6_EARG;
  if (numbrella > 0 && (!kind || !center || !coeffi)) return G6_EARG;
  McUmb u;
  u.numb = numbrella;
  for (int k = 0; k < numbrella; k++) {
    const int t = kind[k];
    if (t != G6_UMB_TRIANGLE && t != G6_UMB_EPOP && t != G6_UMB_LOGTRIAN && t != G6_UMB_Q6LOCAL)
      return G6_EARG;
    u.kind[k] = t;
    u.center[k] = center[k];
    u.coeffi[k] = coeffi[k];
    if (t == G6_UMB_TRIANGLE || t == G6_UMB_LOGTRIAN) u.tri = 1;
    if (t == G6_UMB_Q6LOCAL) u.q6 = 1;
  }
  if (u.tri) {
    if (!refhist) return G6_EARG;
    for (int b = 0; b < 21; b++) u.refhist[b] = refhist[b];
  }
  McBuf &m = s->mc;
  const int n = s->n;
  if (!m.uop) {
    G6_CHECK(hipMalloc(&m.uop, sizeof(double) * kUmbMax));
    G6_CHECK(hipMalloc(&m.uflag, sizeof(int)));
    G6_CHECK(hipMemset(m.uop, 0, sizeof(double) * kUmbMax));
  }
  if (u.q6 && !m.q6.v6) {
    Q6Buf &q = m.q6;
    const size_t nb = (size_t)n * kMaxBond;
    G6_CHECK(hipMalloc(&q.nn6, sizeof(int) * n));
    G6_CHECK(hipMalloc(&q.lab6, sizeof(int) * (size_t)n * kQ6Near));
    G6_CHECK(hipMalloc(&q.v6, sizeof(double) * 4 * (size_t)n * kQ6Near));
    G6_CHECK(hipMalloc(&q.neis_n, sizeof(int) * n));
    G6_CHECK(hipMalloc(&q.neis, sizeof(int) * (size_t)n * kQ6Set));
    G6_CHECK(hipMalloc(&q.neib_n, sizeof(int) * nb));
    G6_CHECK(hipMalloc(&q.neib, sizeof(int) * nb * kQ6Set));
    G6_CHECK(hipMalloc(&q.qlm, sizeof(double) * nb * kQ6Rec));
    G6_CHECK(hipMalloc(&q.qsum, sizeof(double) * (size_t)n * kQ6Rec));
    G6_CHECK(hipMalloc(&q.dq6, sizeof(double) * n));
    G6_CHECK(hipMalloc(&q.op, sizeof(double)));
    G6_CHECK(hipMalloc(&q.nbond, sizeof(int) * n));
    G6_CHECK(hipMalloc(&q.nb3, sizeof(int) * n));
  }
  if (u.tri && !m.tri_v) {
    G6_CHECK(hipMalloc(&m.tri_nn, sizeof(int) * n));
    G6_CHECK(hipMalloc(&m.tri_near, sizeof(int) * (size_t)n * kTriNear));
    G6_CHECK(hipMalloc(&m.tri_hist, sizeof(int) * 21));
    G6_CHECK(hipMalloc(&m.tri_v, sizeof(double) * 4 * (size_t)n * kTriNear));
  }
  s->mc_umb = u;
  return 0;
}

int g6_mc_get_umbrella(g6_handle h, double *op, double *center, double *coeffi) {
  System *s = system_of(h);
  if (!s || !s->mc_ready) return G6_EARG;
  const McUmb &u = s->mc_umb;
  G6_CHECK(hipStreamSynchronize(s->stream));
  if (op && u.numb > 0) G6_CHECK(hipMemcpy(op, s->mc.uop, sizeof(double) * u.numb, hipMemcpyDeviceToHost));
  for (int k = 0; k < u.numb; k++) {
    if (center) center[k] = u.center[k];
    if (coeffi) coeffi[k] = u.coeffi[k];
  }
  return 0;
}

int g6_mc_set_temperature(g6_handle h, double temp) {
  System *s = system_of(h);
  if (!s || !s->mc_ready || !(temp > 0)) return G6_EARG;
  s->mc_beta = 1.0 / temp;
  return 0;
}

int g6_mc_get(g6_handle h, double *com, double *quat1, double *delta, double *rtratio,
              double *epsum) {
  System *s = system_of(h);
  if (!s || !s->mc_ready) return G6_EARG;
  const int n = s->n;
  G6_CHECK(hipStreamSynchronize(s->stream));
  if (com) G6_CHECK(hipMemcpy(com, s->mc.com, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  if (quat1) G6_CHECK(hipMemcpy(quat1, s->mc.quat, sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
  if (delta) G6_CHECK(hipMemcpy(delta, s->mc.delta, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (rtratio) G6_CHECK(hipMemcpy(rtratio, s->mc.rtr, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (epsum) G6_CHECK(hipMemcpy(epsum, s->mc.epsum, sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int g6_mc_pairinfo(g6_handle h, double *ep, double *epsum) {
  System *s = system_of(h);
  if (!s || !s->mc_ready || !ep) return G6_EARG;
  const int n = s->n;
  G6_CHECK(hipStreamSynchronize(s->stream));
  G6_CHECK(hipMemcpy(ep, s->mc.pinfo, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost));
  if (epsum) G6_CHECK(hipMemcpy(epsum, s->mc.epsum, sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

// the total energy of the current MC configuration recomputed from scratch
// (a fresh PairInfo), without touching the bookkeeping
int g6_mc_energy(g6_handle h, double *epsum) {
  System *s = system_of(h);
  if (!s || !s->mc_ready || !epsum) return G6_EARG;
  const int n = s->n;
  double *keep = nullptr;
  G6_CHECK(hipMalloc(&keep, sizeof(double) * (size_t)n * n));
  G6_CHECK(hipMemcpyAsync(keep, s->mc.pinfo, sizeof(double) * (size_t)n * n,
                          hipMemcpyDeviceToDevice, s->stream));
  McArgs a = mc_args(*s);
  k_mc_pairinfo<<<n, 64, 0, s->stream>>>(a, s->std);
  G6_CHECK(hipGetLastError());
  std::vector<double> part(n);
  G6_CHECK(hipMemcpyAsync(part.data(), s->mc.part, sizeof(double) * n, hipMemcpyDeviceToHost,
                          s->stream));
  G6_CHECK(hipMemcpyAsync(s->mc.pinfo, keep, sizeof(double) * (size_t)n * n,
                          hipMemcpyDeviceToDevice, s->stream));
  G6_CHECK(hipStreamSynchronize(s->stream));
  G6_CHECK(hipFree(keep));
  double e = 0.0;
  for (int i = 0; i < n; i++) e = e + part[i];
  *epsum = e;
  return 0;
}

}  // extern "C"
