// NPT Monte Carlo with collective moves, and delayed acceptance screened by
// the learned per-molecule energy surrogate.
//
// The reference has single-molecule NVT moves only (mc2, g6_mc.hip); the
// energy here is the same PairInfo sum (PairInfo.F90:12-46, StdPairPotential
// with CutOff_Dimmer2 on the COM distance, FP64), and two collective moves
// act on it:
//   volume  ln V -> ln V + x, x uniform in [-dlnv, dlnv]; COMs scaled by
//           s = V'^(1/3)/V^(1/3), orientations kept; Metropolis on
//           -beta (dU + P dV) + (N+1) x
//   block   every molecule of one Address_Assign cell (GridModule.F90:503-552,
//           ndiv = int(L/voxa)) displaced uniformly in a cube of half-edge
//           dmax/sqrt(3) and turned by the unit quaternion (1, e) normalised,
//           e uniform in [-rot_e, rot_e]^3; Metropolis on -beta dU
// Proposals come from a counter-based stream (Philox4x32-10 keyed by the
// seed, counter = proposal index and slot), so the device draws them
// itself and a run of moves is one captured graph.  No transcendental
// function enters a proposal: s = exp(x/3) is the degree-5 Taylor
// polynomial (|x/3| < 0.01: truncation below 1e-12 relative), so the CPU
// oracle reproduces every configuration bit for bit.
//
// Plain MC, per move: prep (membership, the moved molecules), eval (one wave
// per changed PairInfo row, FP64), decide (fixed-order sums, Metropolis on
// the device), commit (rows and columns, or the whole matrix).
// Delayed acceptance, per round: K proposals from the same state are drawn
// together; each is screened by the surrogate (descriptors of every molecule
// of the proposed configuration, the MLP on the matrix cores, the change
// against the cached energies of the current configuration); the first one
// that passes the surrogate's Metropolis test goes through the exact
// evaluation above and is accepted with min(1, exp(-beta (dU - dU_s))), so
// the chain samples the exact NPT ensemble.  Proposals after the first
// passing one are discarded unseen (they were drawn independently of
// everything before them, so discarding them changes nothing).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "g6_internal.h"
#include "g6_mlp.h"
#include "g6_std.h"

namespace g6 {

constexpr double kI2Jn = 10.0, kJ2Kn = 0.120273;  // PhysConst.F90:48, 56
constexpr int kNptMaxBlock = 96;   // molecules of one block move
constexpr int kNptSub = 8;         // waves per changed row of a block move
constexpr int kNptMaxK = 64;       // proposals per delayed-acceptance round
constexpr int kDescMaxNb = 128;    // neighbours within the descriptor radius
constexpr float kDescRc = 6.0f;
// The neighbour list of a round: O-O pairs within 6.5 A of the current
// configuration.  A proposal moves an O by less than 0.25 A (block: 0.15 A
// plus the turn of the 0.07 A O offset; volume: 0.03% of 13 A), so every
// pair within 6 A after a proposal is on the list.
constexpr int kNbl = 96;
constexpr double kNblR = 6.5;
constexpr int kCnt = 12;
// counters: 0 proposals, 1 accepted, 2 volume proposals, 3 volume accepted,
// 4 surrogate passes, 5 exact evaluations, 6 rounds, 7 void proposals
// (block larger than kNptMaxBlock), 8 next proposal index

struct NptArgs {
  int n, ns, model, h0, h1, K;
  double rc, rin, c1, voxa;  // mc_switch reads rc, rin, c1
  double beta, press_k;      // 1/T; P/kB in K per A^3
  double pvol, dlnv, dmax, rot_e;
  uint32_t key0, key1;
  double bx[kMaxSite], by[kMaxSite], bz[kMaxSite];
  double *com, *quat, *sx, *sy, *sz, *pinfo;     // current configuration
  double *com2, *sx2, *sy2, *sz2, *pinfo2;       // volume proposal
  double *bc, *bq, *bsx, *bsy, *bsz;             // block proposal, by member slot
  double *rows;                                  // [kNptMaxBlock][n] new rows
  int *members, *slot;                           // [kNptMaxBlock], [n] (-1: not moved)
  double *part;                                  // [n]
  double *st;   // [0] L, [1] U, [2] proposed L, [3] dU, [4] dU_s of the selection
  int *ist;     // [0] members, [1] accepted, [2] selection (-1: none), [3] void, [4] type
  long long *cnt;
  // surrogate screening
  float *fo, *fh, *fL;  // [K][n][3] O, [K][n][6] H positions, [K] box edges
  float *desc, *ys;     // [K*n][32] features, [K*n] energies
  float *ycur;          // [n] energies of the current configuration
  double *des;          // [K] surrogate energy changes
  int *vflag;           // [K] void proposals
  int *derr;            // neighbour overflow of the descriptors
  int *nbl, *nbn;       // [n][kNbl] O-O neighbours within kNblR of the current
                        // configuration (ascending j), [n] their counts
  int *arow, *naff;     // [K][n] molecules whose features a proposal changes, [K] counts
};

// ---- the proposal stream ---------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
__host__ __device__ inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}
// (x + 1/2) / 2^32, exact in FP64, in (0, 1)
__host__ __device__ inline double u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }
constexpr uint32_t kTag = 0x4E505400u;

struct Prop {
  int type;       // 1 volume, 0 block
  double a, x;    // volume: s and x = ln(V'/V); block: a = uniform of the cell choice
  double uacc, u1;  // uniforms of the exact and the surrogate decision
};

// exp(y) for |y| < 0.01 as a fixed polynomial (Horner, no contraction)
__host__ __device__ inline double exp_small(double y) {
  return 1.0 + y * (1.0 + y * (0.5 + y * (1.0 / 6.0 + y * (1.0 / 24.0 + y * (1.0 / 120.0)))));
}

__device__ inline Prop make_prop(const NptArgs &A, unsigned long long idx) {
  uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0u, kTag};
  philox10(c, A.key0, A.key1);
  Prop p;
  p.type = u01(c[0]) < A.pvol ? 1 : 0;
  const double uu = u01(c[1]);
  p.uacc = u01(c[2]);
  p.u1 = u01(c[3]);
  if (p.type) {
    p.x = A.dlnv * (2.0 * uu - 1.0);
    p.a = exp_small(p.x / 3.0);
  } else {
    p.x = 0.0;
    p.a = uu;
  }
  return p;
}

// displacement d and turn qr (unit quaternion) of member slot k
__device__ inline void member_move(const NptArgs &A, unsigned long long idx, int k, double d[3],
                                   double qr[4]) {
  uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)(1 + 2 * k), kTag};
  philox10(c, A.key0, A.key1);
  uint32_t e[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)(2 + 2 * k), kTag};
  philox10(e, A.key0, A.key1);
  for (int a = 0; a < 3; a++) d[a] = A.dmax * (2.0 * u01(c[a]) - 1.0);
  const double e0 = A.rot_e * (2.0 * u01(c[3]) - 1.0), e1 = A.rot_e * (2.0 * u01(e[0]) - 1.0),
               e2 = A.rot_e * (2.0 * u01(e[1]) - 1.0);
  const double r = 1.0 / sqrt(1.0 + e0 * e0 + e1 * e1 + e2 * e2);
  qr[0] = r;
  qr[1] = e0 * r;
  qr[2] = e1 * r;
  qr[3] = e2 * r;
}

// Offset of the block grid of proposal idx, uniform in one cell per axis.
// Cells are those of Address_Assign (GridModule.F90:503-552) shifted by it;
// a random shift per move lets block moves carry molecules across every
// fixed plane while each move keeps its members in their cell.
__device__ inline void block_offset(const NptArgs &A, unsigned long long idx, double L, int nd,
                                    double o[3]) {
  uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0xFFFFu, kTag};
  philox10(c, A.key0, A.key1);
  const double cell = L / nd;
  for (int a = 0; a < 3; a++) o[a] = u01(c[a]) * cell;
}

// Address_Assign cell of a point (GridModule.F90:503-552)
__device__ inline int cell_at(double x, double y, double z, double L, int nd) {
  const double invL = 1.0 / L, cellinv = 1.0 / (L / nd);
  const double c[3] = {x, y, z};
  int jc[3];
  for (int a = 0; a < 3; a++) {
    double d = c[a];
    const double o = round(d * invL) * L;
    d = d - o;
    int j = (int)round(d * cellinv);
    if (j < 0) j += nd;
    j = ((j % nd) + nd) % nd;
    jc[a] = j;
  }
  return (jc[2] * nd + jc[1]) * nd + jc[0];
}

// the cell a block move picks (uniform over the voxa grid of box L)
__device__ inline int block_cell(const NptArgs &A, const Prop &p, double L, int &nd) {
  nd = (int)(L / A.voxa);
  if (nd < 1) nd = 1;
  const int ncell = nd * nd * nd;
  int c = (int)(p.a * (double)ncell);
  if (c > ncell - 1) c = ncell - 1;
  return c;
}

// members of the block of proposal p (cells shifted by o) in index order
// (one block of threads); slot_out[i] = member slot or -1; returns the
// member count
template <typename SlotT>
__device__ int block_members(const NptArgs &A, const Prop &p, double L, const double o[3],
                             int *members_out, SlotT *slot_out, int cap) {
  __shared__ int s_wc[32], s_run;
  const int n = A.n, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int nw = blockDim.x >> 6;
  int nd;
  const int c = block_cell(A, p, L, nd);
  if (tid == 0) s_run = 0;
  __syncthreads();
  for (int base = 0; base < n; base += blockDim.x) {
    const int i = base + tid;
    const bool in = i < n && cell_at(A.com[i] - o[0], A.com[n + i] - o[1],
                                     A.com[2 * n + i] - o[2], L, nd) == c;
    const unsigned long long m = __ballot(in);
    if (lane == 0) s_wc[wid] = __popcll(m);
    __syncthreads();
    int off = s_run, tot = 0;
    for (int w = 0; w < nw; w++) {
      if (w < wid) off += s_wc[w];
      tot += s_wc[w];
    }
    const int pos = off + __popcll(m & ((1ull << lane) - 1ull));
    if (i < n) slot_out[i] = (SlotT)(in ? pos : -1);
    if (in && pos < cap) members_out[pos] = i;
    __syncthreads();
    if (tid == 0) s_run = s_run + tot;
    __syncthreads();
  }
  return s_run;
}

// quaternion product qr (x) q, normalised as qnormalize (QuatModule.F90:215-223)
__device__ inline void turn(const double qr[4], double qa, double qb, double qc, double qd,
                            double out[4]) {
  double w = qr[0] * qa - qr[1] * qb - qr[2] * qc - qr[3] * qd;
  double x = qr[0] * qb + qr[1] * qa + qr[2] * qd - qr[3] * qc;
  double y = qr[0] * qc - qr[1] * qd + qr[2] * qa + qr[3] * qb;
  double z = qr[0] * qd + qr[1] * qc - qr[2] * qb + qr[3] * qa;
  const double r = 1.0 / sqrt(w * w + x * x + y * y + z * z);
  out[0] = w * r;
  out[1] = x * r;
  out[2] = y * r;
  out[3] = z * r;
}

// proposal index of this move (plain) or of the round's selection (DA);
// -1 when the round selected nothing
__device__ inline long long move_index(const NptArgs &A, int da) {
  if (!da) return A.cnt[8];
  const int k = A.ist[2];
  return k < 0 ? -1 : A.cnt[8] + k;
}

// ---- exact path -------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_npt_prep(NptArgs A, int da) {
  const long long idx = move_index(A, da);
  if (idx < 0) return;
  const Prop p = make_prop(A, (unsigned long long)idx);
  const int n = A.n, ns = A.ns, tid = threadIdx.x;
  const double L = A.st[0];
  if (p.type) {
    if (tid == 0) {
      A.st[2] = L * p.a;
      A.ist[0] = 0;
      // one minimum image per pair needs rc < L'/2: a smaller box is void
      A.ist[3] = 2.0 * A.rc >= L * p.a ? 1 : 0;
      A.ist[4] = 1;
    }
    for (int i = tid; i < n; i += blockDim.x) {
      const double cx = A.com[i] * p.a, cy = A.com[n + i] * p.a, cz = A.com[2 * n + i] * p.a;
      A.com2[i] = cx;
      A.com2[n + i] = cy;
      A.com2[2 * n + i] = cz;
      place_sites(A, A.quat[i], A.quat[n + i], A.quat[2 * n + i], A.quat[3 * n + i], cx, cy, cz,
                  A.sx2 + (size_t)i * ns, A.sy2 + (size_t)i * ns, A.sz2 + (size_t)i * ns);
    }
    return;
  }
  int nd;
  const int cell = block_cell(A, p, L, nd);
  double o[3];
  block_offset(A, (unsigned long long)idx, L, nd, o);
  const int cnt = block_members(A, p, L, o, A.members, A.slot, kNptMaxBlock);
  // a member moved out of the cell makes the move void: the reverse move
  // would pick another member set, so q(x'->x) = 0 while q(x->x') > 0
  int out = 0;
  for (int k = tid; cnt <= kNptMaxBlock && k < cnt; k += blockDim.x) {
    const int i = A.members[k];
    double d[3], qr[4], q[4];
    member_move(A, (unsigned long long)idx, k, d, qr);
    turn(qr, A.quat[i], A.quat[n + i], A.quat[2 * n + i], A.quat[3 * n + i], q);
    const double cx = A.com[i] + d[0], cy = A.com[n + i] + d[1], cz = A.com[2 * n + i] + d[2];
    if (cell_at(cx - o[0], cy - o[1], cz - o[2], L, nd) != cell) out = 1;
    A.bc[k * 3] = cx;
    A.bc[k * 3 + 1] = cy;
    A.bc[k * 3 + 2] = cz;
    for (int a = 0; a < 4; a++) A.bq[k * 4 + a] = q[a];
    place_sites(A, q[0], q[1], q[2], q[3], cx, cy, cz, A.bsx + (size_t)k * ns,
                A.bsy + (size_t)k * ns, A.bsz + (size_t)k * ns);
  }
  const int any_out = __syncthreads_or(out);
  if (tid == 0) {
    A.ist[0] = cnt;
    A.ist[3] = (cnt > kNptMaxBlock || any_out) ? 1 : 0;
    A.ist[4] = 0;
  }
}

template <int NS>
__device__ inline void load_sites(const double *x, const double *y, const double *z,
                                  double (&X)[kMaxSite], double (&Y)[kMaxSite],
                                  double (&Z)[kMaxSite]) {
#pragma unroll
  for (int s = 0; s < NS; s++) {
    X[s] = x[s];
    Y[s] = y[s];
    Z[s] = z[s];
  }
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Volume move: row w of the scaled configuration (j > w), written to both
// halves of the proposed matrix, row sum into part[w]
template <int M>
__device__ inline void volume_row(const NptArgs &A, const StdTable &st, int w, int lane) {
  constexpr int NS = StdPattern<M>::NS;
  const int n = A.n;
  double XI[kMaxSite], YI[kMaxSite], ZI[kMaxSite], XJ[kMaxSite], YJ[kMaxSite], ZJ[kMaxSite];
  const double L = A.st[2], invL = 1.0 / L;
  const double xi = A.com2[w], yi = A.com2[n + w], zi = A.com2[2 * n + w];
  load_sites<NS>(A.sx2 + (size_t)w * NS, A.sy2 + (size_t)w * NS, A.sz2 + (size_t)w * NS, XI, YI,
                 ZI);
  double acc = 0.0;
  for (int j = lane; j < n; j += 64) {
    double e = 0.0;
    if (j > w) {
      double d0 = xi - A.com2[j], d1 = yi - A.com2[n + j], d2 = zi - A.com2[2 * n + j];
      const double k0 = round(d0 * invL) * L, k1 = round(d1 * invL) * L, k2 = round(d2 * invL) * L;
      d0 = d0 - k0;
      d1 = d1 - k1;
      d2 = d2 - k2;
      const double dd = d0 * d0 + d1 * d1 + d2 * d2;
      if (dd < A.rc * A.rc) {
        load_sites<NS>(A.sx2 + (size_t)j * NS, A.sy2 + (size_t)j * NS, A.sz2 + (size_t)j * NS, XJ,
                       YJ, ZJ);
        e = std_energy_rr<M>(st, XI, YI, ZI, XJ, YJ, ZJ, k0, k1, k2, mc_switch(A, dd));
      }
      A.pinfo2[(size_t)w * n + j] = e;
      A.pinfo2[(size_t)j * n + w] = e;
      acc = acc + e;
    } else if (j == w) {
      A.pinfo2[(size_t)w * n + w] = 0.0;
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) A.part[w] = acc;
}

// The changed PairInfo rows.  Volume: wave w takes rows w and n-1-w (equal
// work per wave).  Block: kNptSub waves per member row, each a contiguous
// range of j; partial dU of wave (member m, range r) into part[m*kNptSub+r].
template <int M>
__global__ __launch_bounds__(256) void k_npt_eval(NptArgs A, StdTable st, int da) {
  constexpr int NS = StdPattern<M>::NS;
  const long long idx = move_index(A, da);
  if (idx < 0) return;
  const int n = A.n, lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (A.ist[4]) {
    const int h = (n + 1) / 2;
    if (w >= h) return;
    volume_row<M>(A, st, w, lane);
    if (n - 1 - w != w) volume_row<M>(A, st, n - 1 - w, lane);
    return;
  }
  double XI[kMaxSite], YI[kMaxSite], ZI[kMaxSite], XJ[kMaxSite], YJ[kMaxSite], ZJ[kMaxSite];
  const int mslot = w / kNptSub, sub = w % kNptSub;
  if (A.ist[3] || mslot >= A.ist[0]) return;
  const int i = A.members[mslot];
  const int chunk = ((n + kNptSub - 1) / kNptSub + 63) / 64 * 64;
  const int jlo = sub * chunk, jhi = min(n, jlo + chunk);
  const double L = A.st[0], invL = 1.0 / L;
  const double xi = A.bc[mslot * 3], yi = A.bc[mslot * 3 + 1], zi = A.bc[mslot * 3 + 2];
  load_sites<NS>(A.bsx + (size_t)mslot * NS, A.bsy + (size_t)mslot * NS,
                 A.bsz + (size_t)mslot * NS, XI, YI, ZI);
  double acc = 0.0;
  for (int j = jlo + lane; j < jhi; j += 64) {
    if (j == i) {
      A.rows[(size_t)mslot * n + j] = 0.0;
      continue;
    }
    const int sj = A.slot[j];
    double xj, yj, zj;
    if (sj >= 0) {
      xj = A.bc[sj * 3];
      yj = A.bc[sj * 3 + 1];
      zj = A.bc[sj * 3 + 2];
    } else {
      xj = A.com[j];
      yj = A.com[n + j];
      zj = A.com[2 * n + j];
    }
    // PairInfo orientation: the lower index first, d = com_lo - com_hi
    double d0 = i < j ? xi - xj : xj - xi, d1 = i < j ? yi - yj : yj - yi,
           d2 = i < j ? zi - zj : zj - zi;
    const double k0 = round(d0 * invL) * L, k1 = round(d1 * invL) * L, k2 = round(d2 * invL) * L;
    d0 = d0 - k0;
    d1 = d1 - k1;
    d2 = d2 - k2;
    const double dd = d0 * d0 + d1 * d1 + d2 * d2;
    double e = 0.0;
    if (dd < A.rc * A.rc) {
      if (sj >= 0)
        load_sites<NS>(A.bsx + (size_t)sj * NS, A.bsy + (size_t)sj * NS, A.bsz + (size_t)sj * NS,
                       XJ, YJ, ZJ);
      else
        load_sites<NS>(A.sx + (size_t)j * NS, A.sy + (size_t)j * NS, A.sz + (size_t)j * NS, XJ, YJ,
                       ZJ);
      const double er = mc_switch(A, dd);
      e = i < j ? std_energy_rr<M>(st, XI, YI, ZI, XJ, YJ, ZJ, k0, k1, k2, er)
                : std_energy_rr<M>(st, XJ, YJ, ZJ, XI, YI, ZI, k0, k1, k2, er);
    }
    A.rows[(size_t)mslot * n + j] = e;
    if (sj < 0 || j > i) acc = acc + (e - A.pinfo[(size_t)i * n + j]);
  }
  acc = wave_sum(acc);
  if (lane == 0) A.part[w] = acc;
}

// sum of v[0..m) in a fixed order: thread t sums t, t+256, ... then a tree
// over the 256 partials (one block of 256 threads; every thread returns it)
__device__ double tree_sum256(const double *v, int m) {
  __shared__ double s[256];
  double a = 0.0;
  for (int k = threadIdx.x; k < m; k += 256) a = a + v[k];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if (threadIdx.x < h) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + h];
    __syncthreads();
  }
  const double r = s[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(256) void k_npt_decide(NptArgs A, int da) {
  __shared__ int s_acc;
  const long long idx = move_index(A, da);
  if (idx < 0) return;
  const Prop p = make_prop(A, (unsigned long long)idx);
  const int n = A.n, ns = A.ns, tid = threadIdx.x;
  const bool vol = A.ist[4] != 0;
  const bool bad = A.ist[3] != 0;  // void: rejected without a test
  // volume: the new energy (row sums); block: dU (member-range partials)
  const double sum = tree_sum256(A.part, vol ? n : (bad ? 0 : A.ist[0] * kNptSub));
  const double up = sum;
  if (tid == 0) {
    const double L = A.st[0], U = A.st[1];
    const double du = vol ? up - U : (bad ? 0.0 : sum);
    double arg;
    if (da) {
      arg = -((du - A.st[4]) * kI2Jn * kJ2Kn) * A.beta;
    } else if (vol) {
      const double Lp = A.st[2];
      const double dv = Lp * Lp * Lp - L * L * L;
      arg = -(du * kI2Jn * kJ2Kn + A.press_k * dv) * A.beta + (double)(n + 1) * p.x;
    } else {
      arg = -(du * kI2Jn * kJ2Kn) * A.beta;
    }
    const bool acc = !bad && (arg >= 0.0 || p.uacc < exp(arg));
    A.st[3] = du;
    A.ist[1] = acc ? 1 : 0;
    if (!da) {
      A.cnt[0] += 1;
      A.cnt[2] += vol ? 1 : 0;
      A.cnt[7] += bad ? 1 : 0;
    } else {
      A.cnt[5] += 1;
    }
    if (acc) {
      A.cnt[1] += 1;
      A.cnt[3] += vol ? 1 : 0;
      if (vol) {
        A.st[0] = A.st[2];
        A.st[1] = up;
      } else {
        A.st[1] = U + du;
      }
    }
    s_acc = acc ? 1 : 0;
  }
  __syncthreads();
  if (!s_acc || vol) return;
  const int cnt = A.ist[0];
  for (int k = tid; k < cnt; k += 256) {
    const int i = A.members[k];
    for (int a = 0; a < 3; a++) A.com[a * n + i] = A.bc[k * 3 + a];
    for (int a = 0; a < 4; a++) A.quat[a * n + i] = A.bq[k * 4 + a];
    for (int s = 0; s < ns; s++) {
      A.sx[(size_t)i * ns + s] = A.bsx[(size_t)k * ns + s];
      A.sy[(size_t)i * ns + s] = A.bsy[(size_t)k * ns + s];
      A.sz[(size_t)i * ns + s] = A.bsz[(size_t)k * ns + s];
    }
  }
}

// write-back of an accepted move; the last kernel of a move or round, so it
// also advances the proposal index
__global__ __launch_bounds__(256) void k_npt_commit(NptArgs A, int da) {
  const long long idx = move_index(A, da);
  const bool go = idx >= 0 && A.ist[1];
  const int n = A.n, ns = A.ns;
  const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x,
               stride = (size_t)gridDim.x * blockDim.x;
  if (go) {
    if (A.ist[4]) {
      for (size_t t = t0; t < (size_t)n * n; t += stride) A.pinfo[t] = A.pinfo2[t];
      for (size_t t = t0; t < (size_t)3 * n; t += stride) A.com[t] = A.com2[t];
      for (size_t t = t0; t < (size_t)n * ns; t += stride) {
        A.sx[t] = A.sx2[t];
        A.sy[t] = A.sy2[t];
        A.sz[t] = A.sz2[t];
      }
    } else {
      const size_t m = (size_t)A.ist[0] * n;
      for (size_t t = t0; t < m; t += stride) {
        const int w = (int)(t / n), j = (int)(t % n), i = A.members[w];
        const double v = A.rows[t];
        A.pinfo[(size_t)i * n + j] = v;
        A.pinfo[(size_t)j * n + i] = v;
      }
    }
    if (da) {
      const size_t k = (size_t)A.ist[2];
      for (size_t t = t0; t < (size_t)A.naff[k]; t += stride) A.ycur[A.arow[k * n + t]] = A.ys[k * n + t];
    }
  }
  if (t0 == 0) {
    A.cnt[8] += da ? A.K : 1;
    if (go) A.ist[5] = 1;  // the configuration changed: the next round rebuilds the list
  }
}

// ---- surrogate screening ----------------------------------------------------
// O and H sites of configuration k (FP32) and its box edge
__device__ inline void put_sites(const NptArgs &A, int k, int i, const double *x, const double *y,
                                 const double *z) {
  float *o = A.fo + ((size_t)k * A.n + i) * 3;
  float *h = A.fh + ((size_t)k * A.n + i) * 6;
  o[0] = (float)x[0];
  o[1] = (float)y[0];
  o[2] = (float)z[0];
  h[0] = (float)x[A.h0];
  h[1] = (float)y[A.h0];
  h[2] = (float)z[A.h0];
  h[3] = (float)x[A.h1];
  h[4] = (float)y[A.h1];
  h[5] = (float)z[A.h1];
}

// the affected rows of a proposal: every molecule, in index order
__device__ inline void all_rows(const NptArgs &A, int k) {
  for (int i = threadIdx.x; i < A.n; i += blockDim.x) A.arow[(size_t)k * A.n + i] = i;
  if (threadIdx.x == 0) A.naff[k] = A.n;
}

// configuration of proposal k = blockIdx.x of the round, or the current one
// (k = 0, current != 0): FP32 O and H sites, box edge, and the molecules
// whose features it changes (a block: its members and their listed
// neighbours; a volume move: every molecule), in index order
__global__ __launch_bounds__(1024) void k_da_prep(NptArgs A, int current) {
  __shared__ int s_mem[kNptMaxBlock];
  __shared__ short s_slot[4096];
  __shared__ unsigned char s_mark[4096];
  __shared__ int s_wc[16], s_run;
  const int k = blockIdx.x, n = A.n, ns = A.ns, tid = threadIdx.x;
  const double L = A.st[0];
  if (current) {
    for (int i = tid; i < n; i += blockDim.x)
      put_sites(A, 0, i, A.sx + (size_t)i * ns, A.sy + (size_t)i * ns, A.sz + (size_t)i * ns);
    if (tid == 0) A.fL[0] = (float)L;
    all_rows(A, 0);
    return;
  }
  if (k == 0 && tid == 0) A.ist[5] = 0;  // k_nbl of this round has run
  const unsigned long long idx = (unsigned long long)A.cnt[8] + k;
  const Prop p = make_prop(A, idx);
  double X[kMaxSite], Y[kMaxSite], Z[kMaxSite];
  if (p.type) {
    for (int i = tid; i < n; i += blockDim.x) {
      place_sites(A, A.quat[i], A.quat[n + i], A.quat[2 * n + i], A.quat[3 * n + i],
                  A.com[i] * p.a, A.com[n + i] * p.a, A.com[2 * n + i] * p.a, X, Y, Z);
      put_sites(A, k, i, X, Y, Z);
    }
    if (tid == 0) {
      A.fL[k] = (float)(L * p.a);
      A.vflag[k] = 2.0 * A.rc >= L * p.a ? 1 : 0;  // void as in k_npt_prep
    }
    all_rows(A, k);
    return;
  }
  int nd;
  const int cell = block_cell(A, p, L, nd);
  double o[3];
  block_offset(A, idx, L, nd, o);
  const int cnt = block_members(A, p, L, o, s_mem, s_slot, kNptMaxBlock);
  // void as in k_npt_prep: too many members, or a member leaving the cell
  int out = 0;
  for (int m = tid; cnt <= kNptMaxBlock && m < cnt; m += blockDim.x) {
    const int i = s_mem[m];
    double d[3], qr[4];
    member_move(A, idx, m, d, qr);
    if (cell_at(A.com[i] + d[0] - o[0], A.com[n + i] + d[1] - o[1],
                A.com[2 * n + i] + d[2] - o[2], L, nd) != cell)
      out = 1;
  }
  const int any_out = __syncthreads_or(out);
  const bool bad = cnt > kNptMaxBlock || any_out;
  if (tid == 0) {
    A.fL[k] = (float)L;
    A.vflag[k] = bad ? 1 : 0;
  }
  if (bad) {
    if (tid == 0) A.naff[k] = 0;
    return;
  }
  for (int i = tid; i < n; i += blockDim.x) s_mark[i] = 0;
  __syncthreads();
  for (int t = tid; t < cnt * (kNbl + 1); t += blockDim.x) {
    const int m = s_mem[t / (kNbl + 1)], e = t % (kNbl + 1);
    if (e == kNbl) s_mark[m] = 1;
    else if (e < A.nbn[m]) s_mark[A.nbl[(size_t)m * kNbl + e]] = 1;
  }
  for (int i = tid; i < n; i += blockDim.x) {
    const int sl = s_slot[i];
    if (sl < 0) {
      put_sites(A, k, i, A.sx + (size_t)i * ns, A.sy + (size_t)i * ns, A.sz + (size_t)i * ns);
    } else {
      double d[3], qr[4], q[4];
      member_move(A, idx, sl, d, qr);
      turn(qr, A.quat[i], A.quat[n + i], A.quat[2 * n + i], A.quat[3 * n + i], q);
      place_sites(A, q[0], q[1], q[2], q[3], A.com[i] + d[0], A.com[n + i] + d[1],
                  A.com[2 * n + i] + d[2], X, Y, Z);
      put_sites(A, k, i, X, Y, Z);
    }
  }
  // compaction of the marks in index order
  const int lane = tid & 63, wid = tid >> 6, nw = blockDim.x >> 6;
  if (tid == 0) s_run = 0;
  __syncthreads();
  for (int base = 0; base < n; base += blockDim.x) {
    const int i = base + tid;
    const bool in = i < n && s_mark[i];
    const unsigned long long mk = __ballot(in);
    if (lane == 0) s_wc[wid] = __popcll(mk);
    __syncthreads();
    int off = s_run, tot = 0;
    for (int w = 0; w < nw; w++) {
      if (w < wid) off += s_wc[w];
      tot += s_wc[w];
    }
    if (in) A.arow[(size_t)k * n + off + __popcll(mk & ((1ull << lane) - 1ull))] = i;
    __syncthreads();
    if (tid == 0) s_run = s_run + tot;
    __syncthreads();
  }
  if (tid == 0) A.naff[k] = s_run;
}

// O-O neighbours of every molecule within kNblR in the current
// configuration (FP64 minimum image), ascending j; one wave per molecule
__global__ __launch_bounds__(256) void k_nbl(NptArgs A, int always) {
  const int n = A.n, ns = A.ns, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n || !(always || A.ist[5])) return;
  const double L = A.st[0], invL = 1.0 / L;
  const double ox = A.sx[(size_t)i * ns], oy = A.sy[(size_t)i * ns], oz = A.sz[(size_t)i * ns];
  int base = 0;
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int j = j0 + lane;
    bool in = false;
    if (j < n && j != i) {
      double dx = A.sx[(size_t)j * ns] - ox, dy = A.sy[(size_t)j * ns] - oy,
             dz = A.sz[(size_t)j * ns] - oz;
      dx = dx - round(dx * invL) * L;
      dy = dy - round(dy * invL) * L;
      dz = dz - round(dz * invL) * L;
      in = dx * dx + dy * dy + dz * dz < kNblR * kNblR;
    }
    const unsigned long long m = __ballot(in);
    const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
    if (in && pos < kNbl) A.nbl[(size_t)i * kNbl + pos] = j;
    base += __popcll(m);
  }
  if (lane == 0) {
    A.nbn[i] = base < kNbl ? base : kNbl;
    if (base > kNbl) *A.derr = 1;
  }
}

// Per-molecule features of configuration c = wave / n, molecule i = wave % n,
// from neighbours j whose O lies within Rc = 6 A of i's O (FP32):
//   0-7   sum_j exp(-4 (r_OO - mu)^2) fc(r_OO), mu = 2.4, 2.9, ..., 5.9
//   8-11  sum_j sum_{H of i} exp(-10 (r_{H_i O_j} - mu)^2) fc(r_OO), mu = 1.6, 1.9, 2.2, 2.6
//   12-15 the same with the H of j and the O of i
//   16-31 2^(1-z) sum_{j<k} (1 + l cos t_jik)^z exp(-eta (r_ij^2 + r_ik^2)) fc fc,
//         index 16 + 8 e + 4 l + z for eta = 0.02, 0.08; l = +1, -1; z = 1, 2, 4, 8
// with fc(r) = (cos(pi r/Rc) + 1)/2.  Lane-strided sums and a fixed tree: a
// molecule whose neighbourhood is unchanged gets bit-identical features.
__device__ inline float wave_sumf(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_desc(NptArgs A, int nconf) {
  __shared__ float4 s_nb[4][kNbl];
  __shared__ int s_j[4][kNbl];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, n = A.n;
  const long long row = (long long)blockIdx.x * 4 + wid;
  if (row >= (long long)nconf * n) return;
  const int c = (int)(row / n), r = (int)(row % n);
  if (r >= A.naff[c]) return;
  const int i = A.arow[(size_t)c * n + r];
  const float *O = A.fo + (size_t)c * n * 3, *H = A.fh + (size_t)c * n * 6;
  const float L = A.fL[c], iL = 1.0f / L, rc2 = kDescRc * kDescRc, pr = 3.14159265358979f / kDescRc;
  const float ox = O[i * 3], oy = O[i * 3 + 1], oz = O[i * 3 + 2];
  // the listed neighbours within Rc, compacted in list order (ascending j)
  const int nl = A.nbn[i];
  int cnt = 0;
  for (int e0 = 0; e0 < nl; e0 += 64) {
    const int e = e0 + lane;
    float dx = 0, dy = 0, dz = 0, r2 = 1e30f;
    int j = 0;
    if (e < nl) {
      j = A.nbl[(size_t)i * kNbl + e];
      dx = O[j * 3] - ox;
      dy = O[j * 3 + 1] - oy;
      dz = O[j * 3 + 2] - oz;
      dx = dx - L * rintf(dx * iL);
      dy = dy - L * rintf(dy * iL);
      dz = dz - L * rintf(dz * iL);
      r2 = dx * dx + dy * dy + dz * dz;
    }
    const bool in = r2 < rc2;
    const unsigned long long m = __ballot(in);
    const int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
    if (in) {
      const float rr = sqrtf(r2);
      s_nb[wid][pos] = make_float4(dx, dy, dz, 0.5f * (cosf(pr * rr) + 1.0f));
      s_j[wid][pos] = j;
    }
    cnt += __popcll(m);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  float g[32];
#pragma unroll
  for (int k = 0; k < 32; k++) g[k] = 0.0f;
  float hx[2], hy[2], hz[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    hx[h] = H[i * 6 + 3 * h];
    hy[h] = H[i * 6 + 3 * h + 1];
    hz[h] = H[i * 6 + 3 * h + 2];
  }
  // radial and hydrogen-bond terms, one compacted neighbour per lane
  for (int a = lane; a < cnt; a += 64) {
    const float4 v = s_nb[wid][a];
    const int j = s_j[wid][a];
    const float rr = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z), fc = v.w;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const float t = rr - (2.4f + 0.5f * k);
      g[k] += expf(-4.0f * t * t) * fc;
    }
    // the image of j next to i: O_j' = O_i + d, H_j' = H_j + (O_j' - O_j)
    const float sx = ox + v.x - O[j * 3], sy = oy + v.y - O[j * 3 + 1], sz = oz + v.z - O[j * 3 + 2];
    const float mu[4] = {1.6f, 1.9f, 2.2f, 2.6f};
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const float ax = ox + v.x - hx[h], ay = oy + v.y - hy[h], az = oz + v.z - hz[h];
      const float ra = sqrtf(ax * ax + ay * ay + az * az);
      const float bx = H[j * 6 + 3 * h] + sx - ox, by = H[j * 6 + 3 * h + 1] + sy - oy,
                  bz = H[j * 6 + 3 * h + 2] + sz - oz;
      const float rb = sqrtf(bx * bx + by * by + bz * bz);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const float ta = ra - mu[k], tb = rb - mu[k];
        g[8 + k] += expf(-10.0f * ta * ta) * fc;
        g[12 + k] += expf(-10.0f * tb * tb) * fc;
      }
    }
  }
  // angular terms over the pairs a < b of the compacted list, pair p on lane
  // p mod 64 (row a of the triangle holds cnt - 1 - a pairs)
  const int npair = cnt * (cnt - 1) / 2;
  int a = 0, off = 0;  // row a starts at pair index off
  for (int p = lane; p < npair; p += 64) {
    while (p >= off + (cnt - 1 - a)) {
      off += cnt - 1 - a;
      a++;
    }
    const int b = a + 1 + (p - off);
    const float4 va = s_nb[wid][a], vb = s_nb[wid][b];
    const float ra2 = va.x * va.x + va.y * va.y + va.z * va.z;
    const float rb2 = vb.x * vb.x + vb.y * vb.y + vb.z * vb.z;
    const float cs = (va.x * vb.x + va.y * vb.y + va.z * vb.z) / (sqrtf(ra2) * sqrtf(rb2));
    const float f = va.w * vb.w, e2 = ra2 + rb2;
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const float ex = expf(-(e ? 0.08f : 0.02f) * e2) * f;
#pragma unroll
      for (int l = 0; l < 2; l++) {
        const float p1 = l ? 1.0f - cs : 1.0f + cs, p2 = p1 * p1, p4 = p2 * p2, p8 = p4 * p4;
        float *q = g + 16 + 8 * e + 4 * l;
        q[0] += p1 * ex;
        q[1] += p2 * ex;
        q[2] += p4 * ex;
        q[3] += p8 * ex;
      }
    }
  }
  const float zs[4] = {1.0f, 0.5f, 0.125f, 0.0078125f};
#pragma unroll
  for (int k = 0; k < 32; k++) {
    float v = wave_sumf(g[k]);
    if (k >= 16) v *= zs[k & 3];
    g[k] = v;
  }
  if (lane == 0) {
    float *out = A.desc + (size_t)row * 32;
#pragma unroll
    for (int k = 0; k < 32; k++) out[k] = g[k];
  }
}

// Proposal k = blockIdx.x of the round: dU_s, the fixed-order sum over its
// affected rows of (new - cached) surrogate energies, and the surrogate's
// Metropolis test.  vflag[k]: bit 0 void, bit 1 volume, bit 2 passed.
__global__ __launch_bounds__(256) void k_da_sum(NptArgs A) {
  __shared__ double s[256];
  const int n = A.n, k = blockIdx.x, m = A.naff[k];
  double a = 0.0;
  for (int r = threadIdx.x; r < m; r += 256)
    a = a + ((double)A.ys[(size_t)k * n + r] - (double)A.ycur[A.arow[(size_t)k * n + r]]);
  s[threadIdx.x] = a;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if (threadIdx.x < h) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double des = s[0];
  A.des[k] = des;
  const Prop p = make_prop(A, (unsigned long long)A.cnt[8] + k);
  const double L = A.st[0];
  int f = (A.vflag[k] & 1) | (p.type ? 2 : 0);
  if (!(f & 1)) {
    double arg;
    if (p.type) {
      const double Lp = L * p.a;
      const double dv = Lp * Lp * Lp - L * L * L;
      arg = -(des * kI2Jn * kJ2Kn + A.press_k * dv) * A.beta + (double)(n + 1) * p.x;
    } else {
      arg = -(des * kI2Jn * kJ2Kn) * A.beta;
    }
    if (arg >= 0.0 || p.u1 < exp(arg)) f |= 4;
  }
  A.vflag[k] = f;
}

// the round's selection: the first proposal that passed (one thread)
__global__ void k_da_select(NptArgs A) {
  if (threadIdx.x != 0) return;
  int sel = -1, seen = 0, vols = 0, voids = 0;
  for (int k = 0; k < A.K; k++) {
    const int f = A.vflag[k];
    seen++;
    vols += (f >> 1) & 1;
    voids += f & 1;
    if (f & 4) {
      sel = k;
      break;
    }
  }
  A.cnt[0] += seen;
  A.cnt[2] += vols;
  A.cnt[7] += voids;
  A.cnt[4] += sel >= 0 ? 1 : 0;
  A.cnt[9] += (sel >= 0 && (A.vflag[sel] & 2)) ? 1 : 0;  // volume moves among the passes
  A.cnt[6] += 1;
  A.ist[2] = sel;
  A.ist[1] = 0;
  if (sel >= 0) A.st[4] = A.des[sel];
}

// sites of the current configuration
__global__ void k_npt_sites(NptArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const int n = A.n, ns = A.ns;
  place_sites(A, A.quat[i], A.quat[n + i], A.quat[2 * n + i], A.quat[3 * n + i], A.com[i],
              A.com[n + i], A.com[2 * n + i], A.sx + (size_t)i * ns, A.sy + (size_t)i * ns,
              A.sz + (size_t)i * ns);
}

// PairInfo of the current configuration into pinfo (the volume path with
// s = 1 on the current arrays), row sums in part
template <int M>
__global__ __launch_bounds__(256) void k_npt_pairinfo(NptArgs A, StdTable st, double *pinfo) {
  constexpr int NS = StdPattern<M>::NS;
  const int n = A.n, lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= n) return;
  const double L = A.st[0], invL = 1.0 / L;
  double XI[kMaxSite], YI[kMaxSite], ZI[kMaxSite], XJ[kMaxSite], YJ[kMaxSite], ZJ[kMaxSite];
  const double xi = A.com[w], yi = A.com[n + w], zi = A.com[2 * n + w];
  load_sites<NS>(A.sx + (size_t)w * NS, A.sy + (size_t)w * NS, A.sz + (size_t)w * NS, XI, YI, ZI);
  double acc = 0.0;
  for (int j = lane; j < n; j += 64) {
    double e = 0.0;
    if (j > w) {
      double d0 = xi - A.com[j], d1 = yi - A.com[n + j], d2 = zi - A.com[2 * n + j];
      const double k0 = round(d0 * invL) * L, k1 = round(d1 * invL) * L, k2 = round(d2 * invL) * L;
      d0 = d0 - k0;
      d1 = d1 - k1;
      d2 = d2 - k2;
      const double dd = d0 * d0 + d1 * d1 + d2 * d2;
      if (dd < A.rc * A.rc) {
        load_sites<NS>(A.sx + (size_t)j * NS, A.sy + (size_t)j * NS, A.sz + (size_t)j * NS, XJ, YJ,
                       ZJ);
        e = std_energy_rr<M>(st, XI, YI, ZI, XJ, YJ, ZJ, k0, k1, k2, mc_switch(A, dd));
      }
      pinfo[(size_t)w * n + j] = e;
      pinfo[(size_t)j * n + w] = e;
      acc = acc + e;
    } else if (j == w) {
      pinfo[(size_t)w * n + w] = 0.0;
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) A.part[w] = acc;
}

// U = the fixed-order sum of the row sums (the order of k_npt_decide)
__global__ __launch_bounds__(256) void k_npt_total(NptArgs A, double *out) {
  const double u = tree_sum256(A.part, A.n);
  if (threadIdx.x == 0) *out = u;
}

// e_i = (1/2) sum_j pinfo(i, j): one wave per molecule, lane-strided sums
// and a fixed tree
__global__ __launch_bounds__(256) void k_npt_mol_energy(NptArgs A, double *e) {
  const int lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= A.n) return;
  double s = 0.0;
  for (int j = lane; j < A.n; j += 64) s = s + A.pinfo[(size_t)i * A.n + j];
  s = wave_sum(s);
  if (lane == 0) e[i] = 0.5 * s;
}

struct Npt {
  NptArgs a{};
  StdTable st{};
  bool ready = false;
  int kcap = 0;
  int64_t mlp = 0;
  hipGraphExec_t graph = nullptr;
  int graph_moves = 0, graph_mode = -1, graph_k = 0;
  int64_t graph_mlp = 0;
  std::vector<void *> allocs;
  double o_move = 0;  // largest O displacement of a block move (A)
};

static std::mutex g_nmu;
static std::vector<std::pair<System *, Npt *>> g_npt;

static Npt *npt_of(System *s, bool create) {
  std::lock_guard<std::mutex> lk(g_nmu);
  for (auto &e : g_npt)
    if (e.first == s) return e.second;
  if (!create) return nullptr;
  Npt *p = new Npt();
  g_npt.push_back({s, p});
  return p;
}

static void npt_release(Npt &p);

// the NPT state of a system that is being finalised
void npt_forget(System *s) {
  Npt *p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_nmu);
    for (size_t k = 0; k < g_npt.size(); k++)
      if (g_npt[k].first == s) {
        p = g_npt[k].second;
        g_npt.erase(g_npt.begin() + k);
        break;
      }
  }
  if (p) {
    npt_release(*p);
    delete p;
  }
}

static void npt_release(Npt &p) {
  if (p.graph) (void)hipGraphExecDestroy(p.graph);
  p.graph = nullptr;
  for (void *q : p.allocs) (void)hipFree(q);
  p.allocs.clear();
  p.ready = false;
  p.kcap = 0;
}

template <typename T>
static int npt_alloc(Npt &p, T **ptr, size_t count) {
  void *q = nullptr;
  if (hipMalloc(&q, sizeof(T) * (count > 0 ? count : 1)) != hipSuccess) return G6_EHIP;
  p.allocs.push_back(q);
  *ptr = (T *)q;
  return 0;
}
#define NAL(ptr, count)                            \
  do {                                             \
    const int rc_ = npt_alloc(p, &(ptr), (count)); \
    if (rc_) return rc_;                           \
  } while (0)

// the surrogate buffers for K proposals per round
static int npt_alloc_k(Npt &p, int K) {
  if (K <= p.kcap) return 0;
  NptArgs &a = p.a;
  const size_t n = a.n;
  NAL(a.fo, (size_t)K * n * 3);
  NAL(a.fh, (size_t)K * n * 6);
  NAL(a.fL, K);
  NAL(a.desc, (size_t)K * n * 32);
  NAL(a.ys, (size_t)K * n);
  NAL(a.des, K);
  NAL(a.vflag, K);
  NAL(a.arow, (size_t)K * n);
  NAL(a.naff, K);
  p.kcap = K;
  return 0;
}

template <int M>
static void launch_move(System &s, Npt &p, int da) {
  NptArgs &a = p.a;
  k_npt_prep<<<1, 1024, 0, s.stream>>>(a, da);
  const int waves = a.n > kNptMaxBlock * kNptSub ? a.n : kNptMaxBlock * kNptSub;
  k_npt_eval<M><<<(waves + 3) / 4, 256, 0, s.stream>>>(a, p.st, da);
  k_npt_decide<<<1, 256, 0, s.stream>>>(a, da);
  k_npt_commit<<<256, 256, 0, s.stream>>>(a, da);
}

template <int M>
static void launch_round(System &s, Npt &p, Mlp &net) {
  NptArgs &a = p.a;
  k_nbl<<<(a.n + 3) / 4, 256, 0, s.stream>>>(a, 0);
  k_da_prep<<<a.K, 1024, 0, s.stream>>>(a, 0);
  const long long rows = (long long)a.K * a.n;
  k_desc<<<(unsigned)((rows + 3) / 4), 256, 0, s.stream>>>(a, a.K);
  mlp_launch_forward(net, rows, a.desc, nullptr, a.ys, s.stream, a.n, a.naff);
  k_da_sum<<<a.K, 256, 0, s.stream>>>(a);
  k_da_select<<<1, 64, 0, s.stream>>>(a);
  launch_move<M>(s, p, 1);
}

// features of the current configuration into desc rows 0..n-1
static void launch_features(System &s, Npt &p) {
  NptArgs &a = p.a;
  k_nbl<<<(a.n + 3) / 4, 256, 0, s.stream>>>(a, 1);
  k_da_prep<<<1, 1024, 0, s.stream>>>(a, 1);
  k_desc<<<(a.n + 3) / 4, 256, 0, s.stream>>>(a, 1);
}

// surrogate energies of the current configuration into ycur
static void launch_current(System &s, Npt &p, Mlp &net) {
  launch_features(s, p);
  mlp_launch_forward(net, p.a.n, p.a.desc, nullptr, p.a.ycur, s.stream);
}

static int npt_run(System &s, Npt &p, int count, Mlp *net, int K) {
  NptArgs &a = p.a;
  const int da = net ? 1 : 0;
  if (da) {
    const int rc = npt_alloc_k(p, K);
    if (rc) return rc;
    a.K = K;
    launch_current(s, p, *net);
  } else {
    a.K = 1;
  }
  // moves per captured graph (G6_GRAPH_STEPS = 0: eager launches)
  const char *env = getenv("G6_GRAPH_STEPS");
  const int G = env && atoi(env) <= 0 ? 0 : (count < 200 ? count : 200);
  auto body = [&](int m) {
    for (int q = 0; q < m; q++) {
      std_model(s.model, [&](auto mm) {
        constexpr int M = decltype(mm)::value;
        if (!da) launch_move<M>(s, p, 0);
        else launch_round<M>(s, p, *net);
      });
    }
  };
  if (G > 0) {
    const int64_t mid = da ? p.mlp : 0;
    if (!p.graph || p.graph_moves != G || p.graph_mode != da || p.graph_k != a.K ||
        p.graph_mlp != mid) {
      if (p.graph) (void)hipGraphExecDestroy(p.graph);
      p.graph = nullptr;
      hipGraph_t gr;
      G6_CHECK(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
      body(G);
      G6_CHECK(hipStreamEndCapture(s.stream, &gr));
      G6_CHECK(hipGraphInstantiate(&p.graph, gr, nullptr, nullptr, 0));
      G6_CHECK(hipGraphDestroy(gr));
      p.graph_moves = G;
      p.graph_mode = da;
      p.graph_k = a.K;
      p.graph_mlp = mid;
    }
    for (int k = 0; k + G <= count; k += G) G6_CHECK(hipGraphLaunch(p.graph, s.stream));
    body(count % G);
  } else {
    body(count);
  }
  G6_CHECK(hipGetLastError());
  G6_CHECK(hipStreamSynchronize(s.stream));
  return 0;
}

static NptArgs npt_args_base(const System &s) {
  NptArgs a{};
  a.n = s.n;
  a.ns = s.ns;
  a.model = s.model;
  a.h0 = s.model == G6_TIP4P ? 2 : 1;
  a.h1 = s.model == G6_TIP4P ? 3 : 2;
  a.rc = s.rc;
  a.rin = s.rc - s.taper;
  a.c1 = s.taper > 0 ? s.geo.c1 : 0.0;
  for (int k = 0; k < kMaxSite; k++) {
    a.bx[k] = s.mt.bx[k];
    a.by[k] = s.mt.by[k];
    a.bz[k] = s.mt.bz[k];
  }
  return a;
}

}  // namespace g6

using namespace g6;

extern "C" {

int g6_npt_begin(g6_handle h, double temp, double press_bar, uint64_t seed, double disp_max,
                 double rot_max_deg, double dlnv_max, double p_volume, double voxa) {
  System *s = system_of(h);
  if (!s || !(temp > 0) || !(press_bar >= 0) || !(disp_max >= 0) || !(rot_max_deg >= 0) ||
      rot_max_deg >= 90.0 || !(dlnv_max >= 0) || dlnv_max > 0.03 || !(p_volume >= 0) ||
      p_volume > 1 || !(voxa > 0))
    return G6_EARG;
  if (s->n > 4096 || s->n < 2) return G6_EARG;
  if (s->model == G6_OPLSMEOH) return G6_EARG;  // the surrogate's features are water's
  if (!is_cubic(*s)) return G6_EARG;            // volume moves scale a cube
  if (s->rc > 0.5 * s->L * (1.0 - dlnv_max)) return G6_EARG;  // one image within rc
  Npt &p = *npt_of(s, true);
  G6_CHECK(hipStreamSynchronize(s->stream));
  // a second start on the same handle keeps the device buffers
  const bool reuse = p.ready && p.a.n == s->n && p.a.ns == s->ns && p.a.model == s->model;
  if (p.ready && !reuse) npt_release(p);
  NptArgs keep = p.a;
  p.a = npt_args_base(*s);
  NptArgs &a = p.a;
  const size_t n = s->n, ns = s->ns;
  a.voxa = voxa;
  a.beta = 1.0 / temp;
  a.press_k = press_bar * 1e5 * 1e-30 / 1.380066e-23;  // kB of PhysConst.F90
  a.pvol = p_volume;
  a.dlnv = dlnv_max;
  a.dmax = disp_max / sqrt(3.0);
  // max turn 2 atan(|e|) with |e| <= rot_e sqrt(3)
  a.rot_e = tan(0.5 * rot_max_deg * 3.14159265358979323846 / 180.0) / sqrt(3.0);
  a.key0 = (uint32_t)seed;
  a.key1 = (uint32_t)(seed >> 32);
  // an O moves by the displacement plus its turn about the COM, 2 rho sin(theta/2)
  const double rho = sqrt(s->mt.bx[0] * s->mt.bx[0] + s->mt.by[0] * s->mt.by[0] +
                          s->mt.bz[0] * s->mt.bz[0]);
  p.o_move = disp_max + 2.0 * rho * sin(0.5 * rot_max_deg * 3.14159265358979323846 / 180.0);
  if (reuse) {
    const NptArgs fresh = a;
    a = keep;  // every device pointer and the surrogate buffers
    a.voxa = fresh.voxa;
    a.beta = fresh.beta;
    a.press_k = fresh.press_k;
    a.pvol = fresh.pvol;
    a.dlnv = fresh.dlnv;
    a.dmax = fresh.dmax;
    a.rot_e = fresh.rot_e;
    a.key0 = fresh.key0;
    a.key1 = fresh.key1;
    a.K = 1;
    if (p.graph) (void)hipGraphExecDestroy(p.graph);  // captured with the old arguments
    p.graph = nullptr;
  }
  if (!reuse) {
    NAL(a.com, 3 * n);
    NAL(a.quat, 4 * n);
    NAL(a.sx, n * ns);
    NAL(a.sy, n * ns);
    NAL(a.sz, n * ns);
    NAL(a.pinfo, n * n);
    NAL(a.com2, 3 * n);
    NAL(a.sx2, n * ns);
    NAL(a.sy2, n * ns);
    NAL(a.sz2, n * ns);
    NAL(a.pinfo2, n * n);
    NAL(a.bc, 3 * kNptMaxBlock);
    NAL(a.bq, 4 * kNptMaxBlock);
    NAL(a.bsx, kNptMaxBlock * ns);
    NAL(a.bsy, kNptMaxBlock * ns);
    NAL(a.bsz, kNptMaxBlock * ns);
    NAL(a.rows, kNptMaxBlock * n);
    NAL(a.members, kNptMaxBlock);
    NAL(a.slot, n);
    NAL(a.part, n > (size_t)kNptMaxBlock * kNptSub ? n : (size_t)kNptMaxBlock * kNptSub);
    NAL(a.st, 8);
    NAL(a.ist, 8);
    NAL(a.cnt, kCnt);
    NAL(a.ycur, n);
    NAL(a.derr, 1);
    NAL(a.nbl, n * kNbl);
    NAL(a.nbn, n);
    const int rc = npt_alloc_k(p, 1);
    if (rc) return rc;
    a.K = 1;
  }
  make_std_table(s->mt, &p.st);
  bool match = false;
  std_model(s->model, [&](auto m) { match = std_table_matches<decltype(m)::value>(p.st); });
  if (!match) return G6_EARG;
  G6_CHECK(hipMemcpy(a.com, s->md.com, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice));
  G6_CHECK(hipMemcpy(a.quat, s->md.quat, sizeof(double) * 4 * n, hipMemcpyDeviceToDevice));
  double st0[8] = {s->L, 0, 0, 0, 0, 0, 0, 0};
  int ist0[8] = {0, 0, 0, 0, 0, 1, 0, 0};  // [5]: the neighbour list is stale
  long long c0[kCnt] = {0};
  G6_CHECK(hipMemcpy(a.st, st0, sizeof(st0), hipMemcpyHostToDevice));
  G6_CHECK(hipMemcpy(a.ist, ist0, sizeof(ist0), hipMemcpyHostToDevice));
  G6_CHECK(hipMemcpy(a.cnt, c0, sizeof(c0), hipMemcpyHostToDevice));
  G6_CHECK(hipMemset(a.derr, 0, sizeof(int)));
  k_npt_sites<<<(s->n + 255) / 256, 256, 0, s->stream>>>(a);
  std_model(s->model, [&](auto m) {
    k_npt_pairinfo<decltype(m)::value><<<(s->n + 3) / 4, 256, 0, s->stream>>>(a, p.st, a.pinfo);
  });
  k_npt_total<<<1, 256, 0, s->stream>>>(a, a.st + 1);
  G6_CHECK(hipGetLastError());
  G6_CHECK(hipStreamSynchronize(s->stream));
  p.ready = true;
  return 0;
}

int g6_npt_run(g6_handle h, int count, int64_t hm, int K, long long *stats) {
  System *s = system_of(h);
  Npt *p = s ? npt_of(s, false) : nullptr;
  if (!p || !p->ready || count < 0) return G6_EARG;
  Mlp *net = nullptr;
  if (hm) {
    net = mlp_of(hm);
    if (!net || K < 1 || K > kNptMaxK) return G6_EARG;
    // The surrogate is a function of the configuration only while every O-O
    // pair within kDescRc after a proposal is on the round's list of pairs
    // within kNblR before it: a pair closes by at most twice the largest O
    // move of a block, and a volume move scales it by exp(-dlnv/3) at least.
    if (2.0 * p->o_move >= kNblR - (double)kDescRc ||
        kNblR * exp(-p->a.dlnv / 3.0) <= (double)kDescRc)
      return G6_EARG;
    G6_CHECK(hipStreamSynchronize(net->stream));
    p->mlp = hm;
  }
  const int rc = npt_run(*s, *p, count, net, K);
  if (rc) return rc;
  long long c[kCnt];
  int derr = 0;
  G6_CHECK(hipMemcpy(c, p->a.cnt, sizeof(c), hipMemcpyDeviceToHost));
  G6_CHECK(hipMemcpy(&derr, p->a.derr, sizeof(int), hipMemcpyDeviceToHost));
  if (stats)
    for (int k = 0; k < 10; k++) stats[k] = c[k];
  // void proposals (a block of more than kNptMaxBlock members, a member
  // leaving its cell, a box below 2 rc) are rejections: the reverse of a void
  // move is void too.  A descriptor neighbour list that overflowed is an error.
  if (derr) return G6_ECAPACITY;
  return 0;
}

int g6_npt_get(g6_handle h, double *com, double *quat1, double *box_l, double *energy) {
  System *s = system_of(h);
  Npt *p = s ? npt_of(s, false) : nullptr;
  if (!p || !p->ready) return G6_EARG;
  const size_t n = s->n;
  G6_CHECK(hipStreamSynchronize(s->stream));
  double st[8];
  G6_CHECK(hipMemcpy(st, p->a.st, sizeof(st), hipMemcpyDeviceToHost));
  if (com) G6_CHECK(hipMemcpy(com, p->a.com, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  if (quat1) G6_CHECK(hipMemcpy(quat1, p->a.quat, sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
  if (box_l) *box_l = st[0];
  if (energy) *energy = st[1];
  return 0;
}

int g6_npt_energy(g6_handle h, double *energy) {
  System *s = system_of(h);
  Npt *p = s ? npt_of(s, false) : nullptr;
  if (!p || !p->ready || !energy) return G6_EARG;
  NptArgs &a = p->a;
  double *tmp = nullptr;
  G6_CHECK(hipMalloc(&tmp, sizeof(double)));
  // a fresh PairInfo of the current configuration (into the proposal matrix)
  std_model(s->model, [&](auto m) {
    k_npt_pairinfo<decltype(m)::value><<<(s->n + 3) / 4, 256, 0, s->stream>>>(a, p->st, a.pinfo2);
  });
  k_npt_total<<<1, 256, 0, s->stream>>>(a, tmp);
  G6_CHECK(hipGetLastError());
  G6_CHECK(hipMemcpyAsync(energy, tmp, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  G6_CHECK(hipStreamSynchronize(s->stream));
  G6_CHECK(hipFree(tmp));
  return 0;
}

int g6_npt_features(g6_handle h, float *x, double *e) {
  System *s = system_of(h);
  Npt *p = s ? npt_of(s, false) : nullptr;
  if (!p || !p->ready) return G6_EARG;
  NptArgs &a = p->a;
  const size_t n = s->n;
  launch_features(*s, *p);
  double *de = nullptr;
  G6_CHECK(hipMalloc(&de, sizeof(double) * n));
  k_npt_mol_energy<<<(s->n + 3) / 4, 256, 0, s->stream>>>(a, de);
  G6_CHECK(hipGetLastError());
  if (x) G6_CHECK(hipMemcpyAsync(x, a.desc, sizeof(float) * 32 * n, hipMemcpyDeviceToHost, s->stream));
  if (e) G6_CHECK(hipMemcpyAsync(e, de, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
  G6_CHECK(hipStreamSynchronize(s->stream));
  G6_CHECK(hipFree(de));
  int derr = 0;
  G6_CHECK(hipMemcpy(&derr, a.derr, sizeof(int), hipMemcpyDeviceToHost));
  return derr ? G6_ECAPACITY : 0;
}

}  // extern "C"
