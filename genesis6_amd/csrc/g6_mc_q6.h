// The Q6LOCAL_ umbrella of mc2's trial loop on the device: the state of the
// reference's sq6system (Q6Local.F90:42-61) for one loop and its trial steps
// Differentiate (236-465), Integrate (467-700) and Reject (702-750).
//
// The reference updates the per-molecule sums qlmsum incrementally: a trial
// takes the target's old bonds out, puts its new bonds in, and corrects the
// target's own sum for the molecules entering or leaving its 5 A shell; a
// rejected trial adds and subtracts the same terms back.  Floating-point
// additions do not cancel exactly, so the sums carry the history of every
// trial; to give the reference's order parameter they are formed here by the
// same operations in the same order.  The sets of the reference (Set.F90)
// keep their element order (join appends, drop moves the last element into
// the hole, intersection and difference keep the first set's order), and the
// order of the molecules entering or leaving decides the order of the
// corrections to the target's sum, so the sets are kept as ordered lists.
//
// The trial steps run on wave 0 of the trial block with a fixed ownership
// that needs no synchronisation inside the wave: lane c < 26 owns component
// c of every qlm and qlmsum record, lane e owns entry e of every 6 A row,
// lane p & 63 owns position p of every set, and lane 0 owns every count.
// Per-trial lists (the target's new 5 A set, the entering and leaving
// molecules, the new bonds' sets) live in registers the same way.
#pragma once

#include "g6_q6.h"

namespace g6 {

__device__ inline bool q6_is_bond(double dd) { return !(9.0 < dd); }   // radius 3
__device__ inline bool q6_is_shell(double dd) { return !(25.0 < dd); }  // radius 5

// ---- ordered sets in registers ------------------------------------------
struct WSet {
  int s0, s1, n;  // positions lane and lane + 64; n uniform
};
__device__ inline int ws_at(const WSet &s, int p) {
  return __shfl(p < 64 ? s.s0 : s.s1, p & 63, 64);
}
__device__ inline bool ws_has(const WSet &s, int x, int lane) {
  const bool h = (lane < s.n && s.s0 == x) || (lane + 64 < s.n && s.s1 == x);
  return __ballot(h) != 0;
}
__device__ inline void ws_put(WSet &s, int p, int x, int lane) {
  if (lane == (p & 63)) {
    if (p < 64)
      s.s0 = x;
    else
      s.s1 = x;
  }
}
// join (Set.F90): append unless present; *ovf when the list is full
__device__ inline void ws_join(WSet &s, int x, int lane, int &ovf) {
  if (ws_has(s, x, lane)) return;
  if (s.n < kQ6Set) {
    ws_put(s, s.n, x, lane);
    s.n++;
  } else {
    ovf = 1;
  }
}
// drop (Set.F90): the last element moves into the hole
__device__ inline void ws_drop(WSet &s, int x, int lane) {
  const unsigned long long m0 = __ballot(lane < s.n && s.s0 == x);
  const unsigned long long m1 = __ballot(lane + 64 < s.n && s.s1 == x);
  int p;
  if (m0)
    p = __ffsll((long long)m0) - 1;
  else if (m1)
    p = 64 + __ffsll((long long)m1) - 1;
  else
    return;
  const int last = s.n - 1;
  const int v = ws_at(s, last);
  ws_put(s, p, v, lane);
  s.n = last;
}
// intersection (a & b) and difference (a - b) in a's order
__device__ inline WSet ws_and(const WSet &a, const WSet &b, int lane) {
  WSet r{-1, -1, 0};
  for (int p = 0; p < a.n; p++) {
    const int x = ws_at(a, p);
    if (ws_has(b, x, lane)) {
      ws_put(r, r.n, x, lane);
      r.n++;
    }
  }
  return r;
}
__device__ inline WSet ws_minus(const WSet &a, const WSet &b, int lane) {
  WSet r{-1, -1, 0};
  for (int p = 0; p < a.n; p++) {
    const int x = ws_at(a, p);
    if (!ws_has(b, x, lane)) {
      ws_put(r, r.n, x, lane);
      r.n++;
    }
  }
  return r;
}
__device__ inline WSet ws_load(const int *base, const int *cnt, int lane) {
  WSet s;
  s.s0 = base[lane];
  s.s1 = base[lane + 64];
  int c = 0;
  if (lane == 0) c = *cnt;
  s.n = __shfl(c, 0, 64);
  return s;
}
__device__ inline void ws_store(const WSet &s, int *base, int *cnt, int lane) {
  base[lane] = s.s0;
  base[lane + 64] = s.s1;
  if (lane == 0) *cnt = s.n;
}

// ---- 6 A rows in registers ----------------------------------------------
struct WRow {
  int lab, n;  // label of entry lane; n uniform
  double4 v;
};
__device__ inline WRow wr_load(const Q6Buf &q, int i, int lane) {
  WRow r;
  int c = 0;
  if (lane == 0) c = q.nn6[i];
  r.n = __shfl(c, 0, 64);
  r.lab = -1;
  r.v = make_double4(0.0, 0.0, 0.0, 1e300);
  if (lane < r.n) {
    r.lab = q.lab6[(size_t)i * kQ6Near + lane];
    r.v = reinterpret_cast<const double4 *>(q.v6)[(size_t)i * kQ6Near + lane];
  }
  return r;
}
__device__ inline unsigned long long wr_bonds(const WRow &r, int lane) {
  return __ballot(lane < r.n && q6_is_bond(r.v.w));
}
// the number of bond records of a row (at most MAXBOND: a row with more is
// refused before it could be used)
__device__ inline int q6_nbonds(unsigned long long bonds) {
  const int c = __popcll(bonds);
  return c < kMaxBond ? c : kMaxBond;
}
// index of the k-th set bit of m (k < popcount(m))
__device__ inline int q6_select_bit(unsigned long long m, int k) {
  int e = 0;
  for (int b = 32; b >= 1; b >>= 1)
    if (__popcll(m & ((1ull << (e + b)) - 1)) <= k) e += b;
  return e;
}
// sneighbor1_lookup in nei3: the bond index of label x in row r, or -1
__device__ inline int wr_bond_of(const WRow &r, unsigned long long bonds, int x, int lane) {
  const unsigned long long m = __ballot(lane < r.n && r.lab == x) & bonds;
  if (!m) return -1;
  const int e = __ffsll((long long)m) - 1;
  const int jj = __popcll(bonds & ((1ull << e) - 1));
  return jj < kMaxBond ? jj : -1;
}
// label of bond jj of row r
__device__ inline int wr_bond_label(const WRow &r, unsigned long long bonds, int jj) {
  return __shfl(r.lab, q6_select_bit(bonds, jj), 64);
}

// component `lane` (< 26: m = lane / 2 - 6, re or im) of qlm(6, m) of the
// bond with table vector v: the unit vector v / sqrt(dd), then q6.F90's qlm
__device__ inline double q6_component(const double4 &v, int lane) {
  const double pp = 1.0 / sqrt(v.w);
  const double px = v.x * pp, py = v.y * pp, pz = v.z * pp;
  double q[14];
  q6_ylm6(px, py, pz, q);
  const int c = lane < 26 ? lane : 0;
  const int m = c / 2 - 6, am = m < 0 ? -m : m;
  double re = q[2 * am], im = q[2 * am + 1];
  if (m < 0) {
    const double sg = (am & 1) ? -1.0 : 1.0;
    re = sg * re;
    im = sg * -im;
  }
  return (c & 1) ? im : re;
}

// all 26 components (m = -6..6, re, im) of qlm of a bond vector, as above
__device__ inline void q6_record(const double4 &v, double *o) {
  const double pp = 1.0 / sqrt(v.w);
  const double px = v.x * pp, py = v.y * pp, pz = v.z * pp;
  double q[14];
  q6_ylm6(px, py, pz, q);
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

__device__ inline double *q6_rec(const Q6Buf &q, int i, int jj) {
  return q.qlm + ((size_t)i * kMaxBond + jj) * kQ6Rec;
}
__device__ inline int *q6_neib(const Q6Buf &q, int i, int jj) {
  return q.neib + ((size_t)i * kMaxBond + jj) * kQ6Set;
}
__device__ inline int *q6_neib_n(const Q6Buf &q, int i, int jj) {
  return q.neib_n + (size_t)i * kMaxBond + jj;
}

// qlmsum(:, k) +/- qlm * 2, nbond(k) +/- 2 (w = +1 or -1)
__device__ inline void q6_bump(const Q6Buf &q, int k, double qv, int w, int lane) {
  if (lane < 26) {
    double *p = q.qsum + (size_t)k * kQ6Rec + lane;
    *p = w > 0 ? *p + qv * 2.0 : *p - qv * 2.0;
  }
  if (lane == 0) q.nbond[k] += 2 * w;
}

// q6 of one molecule from its sums (Q6Local.F90:429-436)
__device__ inline double q6_of(const Q6Buf &q, int k) {
  const double *s = q.qsum + (size_t)k * kQ6Rec;
  double nq = 0.0;
  for (int m = 0; m < 13; m++) {
    const double a = hypot(s[2 * m], s[2 * m + 1]);
    nq = nq + a * a;
  }
  return sqrt(nq * 4.0 * kPiT / (2.0 * 6 + 1.0)) / (double)q.nbond[k];
}

// What Differentiate leaves for Integrate or Reject: the target's rows
// before and after the move, its new 5 A set, the molecules entering (plus)
// and leaving (minus) it, and the sets and qlm of its new bonds.
struct Q6Trial {
  WRow old_t, new_t;
  unsigned long long old_b, new_b;
  WSet dneis, plus, minus;
  WSet dneib[kMaxBond];
  double dqlm[kMaxBond];
  int nnew;
};

// The bond counts after the move: the target's, and each molecule of its
// row with the bond to the target made or broken.  Returns false when one
// would exceed MAXBOND (the reference's arrays end there).
__device__ inline bool q6_bonds_fit(const Q6Buf &q, const Q6Trial &t, int lane) {
  if (__popcll(t.new_b) > kMaxBond) return false;
  bool bad = false;
  if (lane < t.old_t.n) {
    const int k = t.old_t.lab;
    const int nb = q.nb3[k] - (q6_is_bond(t.old_t.v.w) ? 1 : 0) + (q6_is_bond(t.new_t.v.w) ? 1 : 0);
    bad = nb > kMaxBond;
  }
  return __ballot(bad) == 0;
}

// sq6system_Differentiate up to the q6 recompute: the sums changed in place
__device__ inline void q6_differentiate(const Q6Buf &q, Q6Trial &t, int target, int lane) {
  // 1. the target's old bonds out of every sum of their neighbourhoods
  const int nold = q6_nbonds(t.old_b);
  for (int ii = 0; ii < nold; ii++) {
    const WSet nb = ws_load(q6_neib(q, target, ii), q6_neib_n(q, target, ii), lane);
    const double qv = lane < 26 ? q6_rec(q, target, ii)[lane] : 0.0;
    for (int kk = 0; kk < nb.n; kk++) q6_bump(q, ws_at(nb, kk), qv, -1, lane);
  }
  // 2. the new 5 A set of the target, the molecules entering and leaving it
  const unsigned long long sh = __ballot(lane < t.new_t.n && q6_is_shell(t.new_t.v.w));
  t.dneis = WSet{-1, -1, 0};
  const int ns = __popcll(sh);
  for (int p = 0; p < ns; p++) {
    ws_put(t.dneis, p, __shfl(t.new_t.lab, q6_select_bit(sh, p), 64), lane);
  }
  t.dneis.n = ns;
  int ovf = 0;
  ws_join(t.dneis, target, lane, ovf);
  const WSet neis_t = ws_load(q.neis + (size_t)target * kQ6Set, q.neis_n + target, lane);
  t.plus = ws_minus(t.dneis, neis_t, lane);
  t.minus = ws_minus(neis_t, t.dneis, lane);
  // 3. the new bonds: their sets and qlm, then into the sums
  t.nnew = q6_nbonds(t.new_b);
#pragma unroll
  for (int jj = 0; jj < kMaxBond; jj++) {
    if (jj >= t.nnew) break;
    const int e = q6_select_bit(t.new_b, jj);
    const int j = __shfl(t.new_t.lab, e, 64);
    const WSet neis_j = ws_load(q.neis + (size_t)j * kQ6Set, q.neis_n + j, lane);
    t.dneib[jj] = ws_and(t.dneis, neis_j, lane);
    double4 v;
    v.x = __shfl(t.new_t.v.x, e, 64);
    v.y = __shfl(t.new_t.v.y, e, 64);
    v.z = __shfl(t.new_t.v.z, e, 64);
    v.w = __shfl(t.new_t.v.w, e, 64);
    t.dqlm[jj] = q6_component(v, lane);
  }
#pragma unroll
  for (int jj = 0; jj < kMaxBond; jj++) {
    if (jj >= t.nnew) break;
    for (int kk = 0; kk < t.dneib[jj].n; kk++) q6_bump(q, ws_at(t.dneib[jj], kk), t.dqlm[jj], 1, lane);
  }
  // 4. the target's own sum: bonds of entering molecules now within its
  // shell added, those of leaving ones within its old shell taken out
  for (int ii = 0; ii < t.plus.n; ii++) {
    const int i = ws_at(t.plus, ii);
    const WRow ri = wr_load(q, i, lane);
    const unsigned long long bi = wr_bonds(ri, lane);
    const int nb = q6_nbonds(bi);
    for (int jj = 0; jj < nb; jj++)
      if (ws_has(t.dneis, wr_bond_label(ri, bi, jj), lane))
        q6_bump(q, target, lane < 26 ? q6_rec(q, i, jj)[lane] : 0.0, 1, lane);
  }
  for (int ii = 0; ii < t.minus.n; ii++) {
    const int i = ws_at(t.minus, ii);
    const WRow ri = wr_load(q, i, lane);
    const unsigned long long bi = wr_bonds(ri, lane);
    const int nb = q6_nbonds(bi);
    for (int jj = 0; jj < nb; jj++)
      if (ws_has(neis_t, wr_bond_label(ri, bi, jj), lane))
        q6_bump(q, target, lane < 26 ? q6_rec(q, i, jj)[lane] : 0.0, -1, lane);
  }
}

// sq6system_Reject: Differentiate's terms back, in its order
__device__ inline void q6_reject(const Q6Buf &q, const Q6Trial &t, int target, int lane) {
#pragma unroll
  for (int jj = 0; jj < kMaxBond; jj++) {
    if (jj >= t.nnew) break;
    for (int kk = 0; kk < t.dneib[jj].n; kk++)
      q6_bump(q, ws_at(t.dneib[jj], kk), t.dqlm[jj], -1, lane);
  }
  const int nold = q6_nbonds(t.old_b);
  for (int ii = 0; ii < nold; ii++) {
    const WSet nb = ws_load(q6_neib(q, target, ii), q6_neib_n(q, target, ii), lane);
    const double qv = lane < 26 ? q6_rec(q, target, ii)[lane] : 0.0;
    for (int kk = 0; kk < nb.n; kk++) q6_bump(q, ws_at(nb, kk), qv, 1, lane);
  }
  const WSet neis_t = ws_load(q.neis + (size_t)target * kQ6Set, q.neis_n + target, lane);
  for (int ii = 0; ii < t.plus.n; ii++) {
    const int i = ws_at(t.plus, ii);
    const WRow ri = wr_load(q, i, lane);
    const unsigned long long bi = wr_bonds(ri, lane);
    const int nb = q6_nbonds(bi);
    for (int jj = 0; jj < nb; jj++)
      if (ws_has(t.dneis, wr_bond_label(ri, bi, jj), lane))
        q6_bump(q, target, lane < 26 ? q6_rec(q, i, jj)[lane] : 0.0, -1, lane);
  }
  for (int ii = 0; ii < t.minus.n; ii++) {
    const int i = ws_at(t.minus, ii);
    const WRow ri = wr_load(q, i, lane);
    const unsigned long long bi = wr_bonds(ri, lane);
    const int nb = q6_nbonds(bi);
    for (int jj = 0; jj < nb; jj++)
      if (ws_has(neis_t, wr_bond_label(ri, bi, jj), lane))
        q6_bump(q, target, lane < 26 ? q6_rec(q, i, jj)[lane] : 0.0, 1, lane);
  }
}

// join or drop the target in neib(jj, i) and in the partner's set of the
// same bond, neib(kk, j) with kk the bond of i in j's row
__device__ inline void q6_neib_edit(const Q6Buf &q, int i, int jj, int j, int target, bool join,
                                    int lane, int &ovf) {
  WSet s = ws_load(q6_neib(q, i, jj), q6_neib_n(q, i, jj), lane);
  if (join)
    ws_join(s, target, lane, ovf);
  else
    ws_drop(s, target, lane);
  ws_store(s, q6_neib(q, i, jj), q6_neib_n(q, i, jj), lane);
  const WRow rj = wr_load(q, j, lane);
  const int kk = wr_bond_of(rj, wr_bonds(rj, lane), i, lane);
  if (kk >= 0) {
    WSet u = ws_load(q6_neib(q, j, kk), q6_neib_n(q, j, kk), lane);
    if (join)
      ws_join(u, target, lane, ovf);
    else
      ws_drop(u, target, lane);
    ws_store(u, q6_neib(q, j, kk), q6_neib_n(q, j, kk), lane);
  }
}

// neib(jj, i) = neis(i) & neis(j) for every bond jj of row i
__device__ inline void q6_neib_rebuild(const Q6Buf &q, int i, int lane) {
  const WRow ri = wr_load(q, i, lane);
  const unsigned long long bi = wr_bonds(ri, lane);
  const int nb = q6_nbonds(bi);
  const WSet si = ws_load(q.neis + (size_t)i * kQ6Set, q.neis_n + i, lane);
  for (int jj = 0; jj < nb; jj++) {
    const int j = wr_bond_label(ri, bi, jj);
    const WSet sj = ws_load(q.neis + (size_t)j * kQ6Set, q.neis_n + j, lane);
    const WSet r = ws_and(si, sj, lane);
    ws_store(r, q6_neib(q, i, jj), q6_neib_n(q, i, jj), lane);
  }
}

// sq6system_Integrate (the q6 values themselves are not kept: only their
// sum enters the order parameter)
__device__ inline void q6_integrate(const Q6Buf &q, const Q6Trial &t, int target,
                                    double dx, double dy, double dz, int lane, int &ovf) {
  // the target's qlm records and 5 A set
#pragma unroll
  for (int jj = 0; jj < kMaxBond; jj++) {
    if (jj >= t.nnew) break;
    if (lane < 26) q6_rec(q, target, jj)[lane] = t.dqlm[jj];
  }
  ws_store(t.dneis, q.neis + (size_t)target * kQ6Set, q.neis_n + target, lane);
  for (int ii = 0; ii < t.plus.n; ii++) {
    const int i = ws_at(t.plus, ii);
    WSet s = ws_load(q.neis + (size_t)i * kQ6Set, q.neis_n + i, lane);
    ws_join(s, target, lane, ovf);
    ws_store(s, q.neis + (size_t)i * kQ6Set, q.neis_n + i, lane);
  }
  for (int ii = 0; ii < t.minus.n; ii++) {
    const int i = ws_at(t.minus, ii);
    WSet s = ws_load(q.neis + (size_t)i * kQ6Set, q.neis_n + i, lane);
    ws_drop(s, target, lane);
    ws_store(s, q.neis + (size_t)i * kQ6Set, q.neis_n + i, lane);
  }
  // bonds of the target lost, found and kept (labels in row order)
  WSet oldb{-1, -1, 0}, newb{-1, -1, 0};
  for (int jj = 0; jj < q6_nbonds(t.old_b); jj++) {
    ws_put(oldb, jj, wr_bond_label(t.old_t, t.old_b, jj), lane);
    oldb.n++;
  }
  for (int jj = 0; jj < t.nnew; jj++) {
    ws_put(newb, jj, wr_bond_label(t.new_t, t.new_b, jj), lane);
    newb.n++;
  }
  const WSet lost = ws_minus(oldb, newb, lane), found = ws_minus(newb, oldb, lane),
             kept = ws_and(oldb, newb, lane);
  // the target joins the sets of the entering molecules' bonds whose partner
  // is in its new shell, and leaves those of the leaving molecules' bonds
  // whose partner is in its (now new) shell
  for (int ii = 0; ii < t.plus.n; ii++) {
    const int i = ws_at(t.plus, ii);
    const WRow ri = wr_load(q, i, lane);
    const unsigned long long bi = wr_bonds(ri, lane);
    const int nb = q6_nbonds(bi);
    for (int jj = 0; jj < nb; jj++) {
      const int j = wr_bond_label(ri, bi, jj);
      if (ws_has(t.dneis, j, lane)) q6_neib_edit(q, i, jj, j, target, true, lane, ovf);
    }
  }
  for (int ii = 0; ii < t.minus.n; ii++) {
    const int i = ws_at(t.minus, ii);
    const WRow ri = wr_load(q, i, lane);
    const unsigned long long bi = wr_bonds(ri, lane);
    const int nb = q6_nbonds(bi);
    for (int jj = 0; jj < nb; jj++) {
      const int j = wr_bond_label(ri, bi, jj);
      if (ws_has(t.dneis, j, lane)) q6_neib_edit(q, i, jj, j, target, false, lane, ovf);
    }
  }
  // qlm records of the partners: a lost bond's record drops out, a kept one
  // takes the target's new value, a found one is inserted in label order
  for (int ii = 0; ii < lost.n; ii++) {
    const int i = ws_at(lost, ii);
    const WRow ri = wr_load(q, i, lane);
    const unsigned long long bi = wr_bonds(ri, lane);
    const int nb = q6_nbonds(bi);
    int jj = 0;
    for (int kk = 0; kk < nb; kk++) {
      if (wr_bond_label(ri, bi, kk) != target) {
        if (lane < 26) q6_rec(q, i, jj)[lane] = q6_rec(q, i, kk)[lane];
        jj++;
      }
    }
  }
  double tmp2 = 0.0;
  for (int ii = 0; ii < kept.n; ii++) {
    const int i = ws_at(kept, ii);
    const int jj = wr_bond_of(t.new_t, t.new_b, i, lane);
    if (jj >= 0) {
      for (int b = 0; b < kMaxBond; b++)
        if (b == jj) tmp2 = t.dqlm[b];
    }
    const WRow ri = wr_load(q, i, lane);
    const int kk = wr_bond_of(ri, wr_bonds(ri, lane), target, lane);
    if (kk >= 0 && lane < 26) q6_rec(q, i, kk)[lane] = tmp2;
  }
  for (int ii = 0; ii < found.n; ii++) {
    const int i = ws_at(found, ii);
    // i's row after the move: its entry of the target moved by +delta
    WRow ri = wr_load(q, i, lane);
    if (lane < ri.n && ri.lab == target) {
      ri.v.x = ri.v.x + dx;
      ri.v.y = ri.v.y + dy;
      ri.v.z = ri.v.z + dz;
      ri.v.w = ri.v.x * ri.v.x + ri.v.y * ri.v.y + ri.v.z * ri.v.z;
    }
    const unsigned long long bi = wr_bonds(ri, lane);
    const int m = q6_nbonds(bi);
    double tmp[kMaxBond];
    for (int c = 0; c < kMaxBond; c++)
      tmp[c] = (c + 1 < m && lane < 26) ? q6_rec(q, i, c)[lane] : 0.0;
    const int j0 = wr_bond_of(t.new_t, t.new_b, i, lane);
    if (j0 >= 0) {
      for (int b = 0; b < kMaxBond; b++)
        if (b == j0) tmp2 = t.dqlm[b];
    }
    int jj = 0;
    for (int kk = 0; kk < m && kk < kMaxBond; kk++) {
      double v = tmp2;
      if (wr_bond_label(ri, bi, kk) != target) {
        for (int b = 0; b < kMaxBond; b++)
          if (b == jj) v = tmp[b];
        jj++;
      }
      if (lane < 26) q6_rec(q, i, kk)[lane] = v;
    }
  }
  // the rows: the target's vectors minus the displacement, the others'
  // vector to the target plus it; bond counts with them
  {
    double4 *row = reinterpret_cast<double4 *>(q.v6);
    if (lane < t.new_t.n) row[(size_t)target * kQ6Near + lane] = t.new_t.v;
    if (lane == 0) q.nb3[target] = t.nnew;
    for (int e = 0; e < t.old_t.n; e++) {
      const int k = __shfl(t.old_t.lab, e, 64);
      const WRow rk = wr_load(q, k, lane);
      if (lane < rk.n && rk.lab == target) {
        double4 v = rk.v;
        v.x = v.x + dx;
        v.y = v.y + dy;
        v.z = v.z + dz;
        v.w = v.x * v.x + v.y * v.y + v.z * v.z;
        row[(size_t)k * kQ6Near + lane] = v;
        q.nb3[k] = q.nb3[k] - (q6_is_bond(rk.v.w) ? 1 : 0) + (q6_is_bond(v.w) ? 1 : 0);
      }
    }
  }
  // the bond sets around the target and of the lost partners, afresh
  {
    const WRow rt = wr_load(q, target, lane);
    const unsigned long long bt = wr_bonds(rt, lane);
    const int nb = q6_nbonds(bt);
    for (int jj = 0; jj < nb; jj++) {
      const int j = wr_bond_label(rt, bt, jj);
      const WSet st = ws_load(q.neis + (size_t)target * kQ6Set, q.neis_n + target, lane);
      const WSet sj = ws_load(q.neis + (size_t)j * kQ6Set, q.neis_n + j, lane);
      ws_store(ws_and(st, sj, lane), q6_neib(q, target, jj), q6_neib_n(q, target, jj), lane);
      q6_neib_rebuild(q, j, lane);
    }
  }
  for (int ii = 0; ii < lost.n; ii++) q6_neib_rebuild(q, ws_at(lost, ii), lane);
}

}  // namespace g6
