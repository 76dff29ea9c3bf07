// The FP32 MD force kernel on cell tiles: every unordered molecule pair is
// evaluated once (Newton's third law), with the tile's neighbourhood staged in
// LDS.
//
// Grid.  The MD grid of a tile system has cells of edge >= rc + skin, so every
// pair within rc + skin lies in one cell or in two cells one step apart per
// axis (ndiv >= 3 keeps the 27 offsets distinct).  A tile is one cell; its
// neighbourhood is the cell itself (direction 0) and its 13 forward cells
// (offsets o with oz > 0, or oz = 0 and oy > 0, or oz = oy = 0 and ox > 0),
// each cell's residents sorted by molecule id.  A pair in one cell belongs to
// the row of its lower id; a pair in two cells to the row of the molecule
// that sees the other's cell in a forward direction.  Every pair within
// rc + skin is therefore in exactly one half row: the reference's half-shell
// stencil (NeighborCell_ListHomo, GridModule.F90:279-465) applied to tiles.
//
// One workgroup per (tile, slice); slice s of S takes the members k with
// k mod S = s.  Per step:
//   1. neighbourhood ids: on a row-build step from the cell lists of the
//      step, else as frozen at the last build;
//   2. staging: the FP64 COM (and, with STAGE, the FP32 site offsets) of every
//      slot into LDS as component arrays; six 64-bit fixed-point force and
//      torque accumulators per slot zeroed;
//   3. per wave and member: on a build step the half row (FP64 minimum image,
//      dd < (rc+skin)^2, in slot order, stored); then the reference's FP64
//      interaction test dd < rc^2 (Interaction_Compress2,
//      InteractionModule.F90:398-605) on each entry.  Passing pairs go into a
//      ring in LDS, each member's run padded to a multiple of four; every full
//      round of 64 is evaluated densely with pair_on_i (FP32).  The quad's sum
//      of the i-side force and torque, and each lane's j-side (F_j = -F_i,
//      tau_j = -tau_i - d x F_i by the pair's angular momentum balance), are
//      added to the fixed-point accumulators with 64-bit LDS atomics.  Integer
//      sums do not depend on the order of the additions;
//   4. every slot's accumulators as FP32 partials part[m][d*S + s][6];
//   5. energy and pair virial: FP64 lane sums, compensated wave and block
//      sums, one partial per workgroup for the step close.
// The corrector (k_md_pre) sums each molecule's partials in a fixed order, so
// the step is deterministic.
#include "g6_md_common.h"

namespace g6 {

constexpr double kFix = 16777216.0;  // 2^24, scale of the fixed-point accumulators
constexpr int kSlotBits = 11;        // row entry: slot | image code << 11
constexpr int kSlotMask = (1 << kSlotBits) - 1;
// ring entries per wave: a member's chunk of 64 candidates is pushed onto
// fewer than 64 unevaluated entries, and full rounds are evaluated after
// every push or pad, so at most 127 entries are ever pending
constexpr int kRing = 128;
constexpr unsigned kNoSlot = 0xFFFFu;
constexpr int kRowBuf = 256;         // row entries of the next member prefetched per wave

// forward direction d = 1..13 of the 3^3 stencil: the offsets after the
// centre in the enumeration e = (oz+1)*9 + (oy+1)*3 + (ox+1)
__device__ inline void tile_dir(int d, int &ox, int &oy, int &oz) {
  const int e = 13 + d;
  ox = e % 3 - 1;
  oy = (e / 3) % 3 - 1;
  oz = e / 9 - 1;
}

// round(v * 2^24) as a two's-complement 64-bit integer (exact rounding for
// |v| < 2^27): the low mantissa bits of v*2^24 + 1.5*2^52
__device__ inline unsigned long long fix64(float v) {
  const double m = 6755399441055744.0;
  const double t = __builtin_fma((double)v, kFix, m);
  return __builtin_bit_cast(unsigned long long, t) - __builtin_bit_cast(unsigned long long, m);
}
__device__ inline float unfix(unsigned long long a) {
  return (float)((double)(long long)a * (1.0 / kFix));
}

// sum over the four lanes of a quad; every lane gets the same bits
__device__ inline float quad_sum(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF,
                                                           0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x4E, 0xF,
                                                           0xF, false));
  return v;
}

__device__ inline double readlane_d(double v, int l) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// compensated sum over the 64 lanes: rows of 16 (TwoSum), then the four row
// totals in lane order; the result is wave-uniform
__device__ inline void wsum_dd(double x, double &s, double &c) {
  double a = x, e = 0.0;
  rowsum16_dd(a, e);
  s = readlane_d(a, 0);
  c = readlane_d(e, 0);
#pragma unroll
  for (int r = 1; r < 4; r++) dd_add(s, c, readlane_d(a, 16 * r), readlane_d(e, 16 * r));
}

template <int M>
__device__ inline void sites_lds(const float *s_site, int nbmax, int q,
                                 float (&r)[Traits<M>::NS][3]) {
#pragma unroll
  for (int s = 0; s < Traits<M>::COM; s++)
#pragma unroll
    for (int a = 0; a < 3; a++) r[s][a] = s_site[(s * 3 + a) * nbmax + q];
#pragma unroll
  for (int a = 0; a < 3; a++) r[Traits<M>::COM][a] = 0.0f;
}

template <int M, bool STAGE>
__device__ __forceinline__ void tile_body(const Geom &g, const MdBuf &b, const PairK<float> &pk,
                                          int parity, const TileLayout &lay, int blk_hw, int nblk) {
  // Workgroups are dealt to the eight XCDs round-robin by index.  With xcd,
  // the workgroups of one XCD (index mod 8) take one contiguous run of
  // tiles, a slab of the box in z-major order, so the records its tiles
  // share stay in that XCD's L2.
  int blk = blk_hw;
  if (lay.xcd) {
    const int q = nblk >> 3, r = nblk & 7, x = blk_hw & 7, p = blk_hw >> 3;
    blk = x < r ? x * (q + 1) + p : r * (q + 1) + (x - r) * q + p;
  }
  using T = Traits<M>;
  constexpr int NQ = T::COM;
  extern __shared__ __align__(16) unsigned char s_dyn[];
  __shared__ int s_base[16], s_cell[14], s_ccnt[14];
  __shared__ int s_wcnt[16][2];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int W = blockDim.x >> 6;
  const int S = b.slices;
  const int tile = blk / S, sl = blk - tile * S;
  const int nd = g.nd[0], n = g.n;  // tiles run cubic boxes
  const int tx = tile % nd, ty = (tile / nd) % nd, tz = tile / (nd * nd);
  const int nbmax = lay.nbmax;
  int *s_id = reinterpret_cast<int *>(s_dyn + lay.id);
  double *s_x = reinterpret_cast<double *>(s_dyn + lay.cx);
  double *s_y = reinterpret_cast<double *>(s_dyn + lay.cy);
  double *s_z = reinterpret_cast<double *>(s_dyn + lay.cz);
  float *s_site = reinterpret_cast<float *>(s_dyn + lay.site);
  unsigned long long *s_acc = reinterpret_cast<unsigned long long *>(s_dyn + lay.acc);
  unsigned *r_ij = reinterpret_cast<unsigned *>(s_dyn + lay.ring + (size_t)wid * 16 * kRing);
  float *r_d = reinterpret_cast<float *>(r_ij + kRing);  // [3][kRing]
  unsigned *r_row = reinterpret_cast<unsigned *>(s_dyn + lay.rowbuf + (size_t)wid * 4 * kRowBuf);
  int *s_order = reinterpret_cast<int *>(s_dyn + lay.order);       // members' work order
  int *s_mc = reinterpret_cast<int *>(s_dyn + lay.ring);           // their row lengths (prologue)
  const int rebuild = b.flags[2 + parity];

  // ---- 1. neighbourhood ids ----
  if (rebuild) {
    if (tid < 14) {
      int ox = 0, oy = 0, oz = 0;
      if (tid > 0) tile_dir(tid, ox, oy, oz);
      const int c = (((tz + oz + nd) % nd) * nd + (ty + oy + nd) % nd) * nd + (tx + ox + nd) % nd;
      s_cell[tid] = c;
      s_ccnt[tid] = min(b.cellres[((size_t)parity * g.ncell + c) * b.cell_cap], b.cell_cap - 1);
    }
    __syncthreads();
    if (tid == 0) {
      int acc = 0;
      for (int d = 0; d < 14; d++) {
        s_base[d] = acc;
        acc += s_ccnt[d];
      }
      if (acc > nbmax) raise_flag(b.flags, 4);  // neighbourhood beyond the LDS slots
      s_base[14] = min(acc, nbmax);
    }
    __syncthreads();
    // each cell's residents, sorted by id, into its slots (wave w: cells w, w+W, ...)
    for (int d = wid; d < 14; d += W) {
      const int cnt = s_ccnt[d];
      const int *res = b.cellres + ((size_t)parity * g.ncell + s_cell[d]) * b.cell_cap + 1;
      int *tmp = reinterpret_cast<int *>(r_ij);  // 4*kRing ints of wave scratch
      for (int r = lane; r < cnt; r += 64) tmp[r] = res[r];
      wsync();
      for (int r = lane; r < cnt; r += 64) {
        const int key = tmp[r];
        int rank = 0;
        for (int u = 0; u < cnt; u++) rank += tmp[u] < key;
        const int slot = s_base[d] + rank;
        if (slot < nbmax) s_id[slot] = key;
      }
      wsync();
    }
    __syncthreads();
    if (sl == 0) {  // frozen until the next build
      for (int q = tid; q < s_base[14]; q += blockDim.x)
        b.nb_id[(size_t)tile * nbmax + q] = s_id[q];
      if (tid < 15) b.nb_base[tile * 16 + tid] = s_base[tid];
    }
  } else {
    // the ids and the members' row lengths are fetched with the slot bases
    // (entries past the counts unused)
    if (tid < 15) s_base[tid] = b.nb_base[tile * 16 + tid];
    for (int q = tid; q < nbmax; q += blockDim.x) s_id[q] = b.nb_id[(size_t)tile * nbmax + q];
    for (int q = tid; q < b.cell_cap; q += blockDim.x)
      s_mc[q] = b.tcnt[(size_t)tile * b.cell_cap + q];
  }
  __syncthreads();
  const int nn = s_base[14];
  const int n0 = min(s_base[1], nn);  // the tile's own molecules: slots 0..n0-1
  // Work order of the slice's members k = sl + S*m: by row length, longest
  // first (ties by m), dealt to the waves in snake order (wave w takes ranks
  // w, 2W-1-w, 2W+w, ...), which balances the waves' pair counts.  Rows are
  // rebuilt on a build step, whose members keep their index order.
  const int Ms = n0 > sl ? (n0 - sl + S - 1) / S : 0;
  for (int t = tid; t < Ms; t += blockDim.x) {
    const int c = rebuild ? 0 : s_mc[sl + S * t];
    int r = 0;
    for (int u = 0; u < Ms; u++) {
      const int cu = rebuild ? 0 : s_mc[sl + S * u];
      r += (cu > c) || (cu == c && u < t);
    }
    s_order[r] = t;
  }
  __syncthreads();
  // the t-th member of this wave, or -1
  auto member_at = [&](int t) -> int {
    const int r = t * W + ((t & 1) ? (W - 1 - wid) : wid);
    return r < Ms ? sl + S * s_order[r] : -1;
  };
  // the row of the next member is fetched into registers while the current
  // one is processed, then parked in the wave's LDS row buffer
  unsigned nx[kRowBuf / 64] = {0u, 0u, 0u, 0u};
  int ncnt = 0;
  auto fetch = [&](int kn) {
    if (rebuild || kn < 0) return;
    const int in = s_id[kn];
    ncnt = b.trow_cnt[in];
    const unsigned *rw = b.trow + (size_t)in * b.maxrow;
#pragma unroll
    for (int u = 0; u < kRowBuf / 64; u++) {
      const int c = 64 * u + lane;
      nx[u] = c < b.maxrow ? rw[c] : 0u;
    }
  };
  auto park = [&]() {
#pragma unroll
    for (int u = 0; u < kRowBuf / 64; u++) r_row[64 * u + lane] = nx[u];
  };
  fetch(member_at(0));  // in flight during the staging

  // ---- 2. staging ----
  for (int q = tid; q < nn; q += blockDim.x) {
    const double *rec = b.rec + (size_t)s_id[q] * b.rs;
    const double2 c01 = *reinterpret_cast<const double2 *>(rec);
    s_x[q] = c01.x;
    s_y[q] = c01.y;
    s_z[q] = rec[2];
    if (STAGE) {
      constexpr int NV = (NQ * 3 + 3) / 4;
      const float4 *p = reinterpret_cast<const float4 *>(rec + 4);
      float f[NV * 4];
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const float4 t = p[v];
        f[4 * v] = t.x; f[4 * v + 1] = t.y; f[4 * v + 2] = t.z; f[4 * v + 3] = t.w;
      }
#pragma unroll
      for (int e = 0; e < NQ * 3; e++) s_site[e * nbmax + q] = f[e];
    }
#pragma unroll
    for (int c = 0; c < 6; c++) s_acc[c * nbmax + q] = 0ull;
  }
  __syncthreads();

  park();

  // ---- 3. rows, interaction test, dense pair rounds ----
  const double Lb = g.L[0], invL = g.invL[0], rr = g.rr, rc2 = g.rc2;
  const double fmax = 33554432.0 / (1.0 + sqrt(rc2));
  const float m2max = (float)(fmax * fmax);
  // each lane's energy and virial over its own pairs (a few per step) in
  // FP32; the lane totals go to FP64 for the compensated block sums
  float ep = 0.0f, v[9];
#pragma unroll
  for (int q = 0; q < 9; q++) v[q] = 0.0f;
  int npair = 0, nlist = 0;  // wave-uniform
  int wp = 0, rp = 0;        // ring write / read positions (wave-uniform)
  const unsigned long long lt = (1ull << lane) - 1ull;

  // one round of 64 ring entries from rp; lanes at or past `live` are idle
  auto round64 = [&](int live) {
    wsync();
    const int pos = (rp & (kRing - 1)) + lane;
    const bool on = lane < live;
    const unsigned e = on ? r_ij[pos] : kNoSlot << 16;
    const int k = (int)(e & 0xFFFFu);
    const unsigned jslot = e >> 16;
    const bool valid = on && jslot != kNoSlot;
    float d0 = 1000.0f, d1 = 0.0f, d2 = 0.0f;
    if (on) {
      d0 = r_d[pos];
      d1 = r_d[kRing + pos];
      d2 = r_d[2 * kRing + pos];
    }
    const int js = valid ? (int)jslot : k;
    float ri[T::NS][3], rj[T::NS][3];
    if (STAGE) {
      sites_lds<M>(s_site, nbmax, k, ri);
      sites_lds<M>(s_site, nbmax, js, rj);
    } else {
      RecSites<M, float>::load(b.rec + (size_t)s_id[k] * b.rs, ri);
      RecSites<M, float>::load(b.rec + (size_t)s_id[js] * b.rs, rj);
    }
    float F[3], Tq[3];
    const float e1 = pair_on_i<M, float>(pk, d0 * d0 + d1 * d1 + d2 * d2, d0, d1, d2, ri, rj, F, Tq);
    if (!valid) {
#pragma unroll
      for (int a = 0; a < 3; a++) F[a] = Tq[a] = 0.0f;
    }
    // i side: the quad holds one member
    float Fq[3], Tqq[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      Fq[a] = quad_sum(F[a]);
      Tqq[a] = quad_sum(Tq[a]);
    }
    if (on && (lane & 3) == 0) {
#pragma unroll
      for (int a = 0; a < 3; a++) {
        atomicAdd(&s_acc[a * nbmax + k], fix64(Fq[a]));
        atomicAdd(&s_acc[(3 + a) * nbmax + k], fix64(Tqq[a]));
      }
    }
    if (valid) {
      const float tj[3] = {-(Tq[0] + (d1 * F[2] - d2 * F[1])),
                           -(Tq[1] + (d2 * F[0] - d0 * F[2])),
                           -(Tq[2] + (d0 * F[1] - d1 * F[0]))};
      // fix64 rounds exactly below 2^27.  With |F|, |tau_i| below
      // 2^25 / (1 + rc) (the norm test) and |d| < rc, every summed value
      // (tau_j, and each quad's sum of four) stays below 2^27; a larger pair
      // force (molecules overlapping in an unphysical configuration) is a
      // capacity error
      const float m2 = F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + Tq[0] * Tq[0] +
                       Tq[1] * Tq[1] + Tq[2] * Tq[2];
      if (!(m2 < m2max)) raise_flag(b.flags, 6);
#pragma unroll
      for (int a = 0; a < 3; a++) {
        atomicAdd(&s_acc[a * nbmax + js], fix64(-F[a]));
        atomicAdd(&s_acc[(3 + a) * nbmax + js], fix64(tj[a]));
      }
      ep += e1;
      const float dd3[3] = {d0, d1, d2};
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) v[3 * a + c] = __builtin_fmaf(dd3[a], F[c], v[3 * a + c]);
    }
    rp += 64;
  };

  // The wave's work as one loop with a single evaluation site: a pass either
  // evaluates a round (64 entries pending, or the last entries once the
  // members are done), or starts a member, pushes the next 64 of its
  // candidates, or finishes it (build bookkeeping, padding, the next row
  // parked).  Pushes happen with fewer than 64 entries pending, so the ring
  // never holds more than 127.
  int t = 0, k = member_at(0), i = 0, c0 = 0, ncand = 0, base = 0;
  bool open = false;
  double xi = 0.0, yi = 0.0, zi = 0.0;
  for (;;) {
    const int pend = wp - rp;
    if (pend >= 64 || (k < 0 && pend > 0)) {
      round64(pend >= 64 ? 64 : pend);
      continue;
    }
    if (k < 0) break;
    if (!open) {
      xi = s_x[k];
      yi = s_y[k];
      zi = s_z[k];
      i = s_id[k];
      // candidates: the slots after k (the same cell with a larger id, then
      // the forward cells) on a build step, else the stored row
      ncand = rebuild ? nn - k - 1 : ncnt;
      fetch(member_at(t + 1));
      c0 = 0;
      base = 0;  // row length so far (build step)
      open = true;
      wsync();   // the parked row
      continue;
    }
    if (c0 < ncand) {
      unsigned *row = b.trow + (size_t)i * b.maxrow;
      const int c = c0 + lane;
      bool pass = false;
      int q = 0;
      double d0 = 0.0, d1 = 0.0, d2 = 0.0;
      if (rebuild) {
        bool keep = false;
        unsigned ent = 0;
        if (c < ncand) {
          q = k + 1 + c;
          d0 = xi - s_x[q];
          d1 = yi - s_y[q];
          d2 = zi - s_z[q];
          const double k0 = round(d0 * invL), k1 = round(d1 * invL), k2 = round(d2 * invL);
          d0 = d0 - k0 * Lb;
          d1 = d1 - k1 * Lb;
          d2 = d2 - k2 * Lb;
          const double dd = d0 * d0 + d1 * d1 + d2 * d2;
          keep = dd < rr;
          if (keep && (fabs(k0) > 63.0 || fabs(k1) > 63.0 || fabs(k2) > 63.0)) {
            raise_flag(b.flags, 2);
            keep = false;
          }
          const unsigned code = (unsigned)(((int)k0 + 64) | (((int)k1 + 64) << 7) |
                                           (((int)k2 + 64) << 14));
          ent = (unsigned)q | (code << kSlotBits);
          pass = keep && dd < rc2;
        }
        const unsigned long long km = __ballot(keep);
        if (keep) {
          const int p = base + __popcll(km & lt);
          if (p < b.maxrow) row[p] = ent;
          else raise_flag(b.flags, 4);  // row capacity exceeded
        }
        base += __popcll(km);
      } else if (c < ncand) {
        const unsigned ent = c < kRowBuf ? r_row[c] : row[c];
        q = (int)(ent & kSlotMask);
        const unsigned code = ent >> kSlotBits;
        const double ox = (double)((int)(code & 127) - 64) * Lb;
        const double oy = (double)((int)((code >> 7) & 127) - 64) * Lb;
        const double oz = (double)((int)((code >> 14) & 127) - 64) * Lb;
        d0 = (xi - s_x[q]) - ox;
        d1 = (yi - s_y[q]) - oy;
        d2 = (zi - s_z[q]) - oz;
        pass = d0 * d0 + d1 * d1 + d2 * d2 < rc2;
      }
      const unsigned long long pm = __ballot(pass);
      if (pass) {
        const int p = (wp + __popcll(pm & lt)) & (kRing - 1);
        r_ij[p] = (unsigned)k | ((unsigned)q << 16);
        r_d[p] = (float)d0;
        r_d[kRing + p] = (float)d1;
        r_d[2 * kRing + p] = (float)d2;
      }
      wp += __popcll(pm);
      npair += __popcll(pm);
      c0 += 64;
      continue;
    }
    // the member is done
    if (rebuild) {
      const int m = min(base, b.maxrow);
      nlist += m;
      if (lane == 0) {
        b.trow_cnt[i] = m;
        b.tcnt[(size_t)tile * b.cell_cap + k] = m;
        b.com_ref[i] = xi;
        b.com_ref[n + i] = yi;
        b.com_ref[2 * n + i] = zi;
      }
    }
    // pad the member's run to a multiple of four
    const int pad = (-wp) & 3;
    if (lane < pad) {
      const int p = (wp + lane) & (kRing - 1);
      r_ij[p] = (unsigned)k | (kNoSlot << 16);
      r_d[p] = 1000.0f;
      r_d[kRing + p] = 0.0f;
      r_d[2 * kRing + p] = 0.0f;
    }
    wp += pad;
    wsync();  // every lane has read the parked row
    if (!rebuild) park();
    t++;
    k = member_at(t);
    open = false;
  }
  if (lane == 0) {
    s_wcnt[wid][0] = npair;
    s_wcnt[wid][1] = nlist;
  }
  __syncthreads();

  // ---- 4. partials of every slot ----
  for (int q = tid; q < nn; q += blockDim.x) {
    int d = 0;
    while (d < 13 && q >= s_base[d + 1]) d++;
    float2 *p = reinterpret_cast<float2 *>(b.part + ((size_t)s_id[q] * b.nds + d * S + sl) * 6);
    p[0] = make_float2(unfix(s_acc[q]), unfix(s_acc[nbmax + q]));
    p[1] = make_float2(unfix(s_acc[2 * nbmax + q]), unfix(s_acc[3 * nbmax + q]));
    p[2] = make_float2(unfix(s_acc[4 * nbmax + q]), unfix(s_acc[5 * nbmax + q]));
  }
  __syncthreads();

  // ---- 5. energy and virial: the lane partials through LDS (the
  // accumulators' region), then compensated block sums in a fixed order:
  // thread (q, seg) takes lanes seg, seg+16, ... of quantity q (TwoSum), and
  // the 16 segments of q, one DPP row, are combined with compensated adds
  const int nl = 64 * W;
  double *red = reinterpret_cast<double *>(s_dyn + lay.acc);  // [10][64 W]
  red[tid] = (double)ep;
#pragma unroll
  for (int q = 0; q < 9; q++) red[(1 + q) * nl + tid] = (double)v[q];
  __syncthreads();
  if (tid < 160) {
    const int q = tid >> 4, seg = tid & 15;
    double r = 0.0, rcmp = 0.0;
    for (int u = seg; u < nl; u += 16) two_sum_acc(r, rcmp, red[q * nl + u]);
    rowsum16_dd(r, rcmp);
    if (seg == 0) b.blockpart[((size_t)parity * nblk + blk) * kBP + q] = r + rcmp;
  } else if (tid < 162) {
    // pair counts: [10] interacting pairs, [11] row entries of a build step
    const int q = tid - 160;
    int t = 0;
    for (int w = 0; w < W; w++) t += s_wcnt[w][q];
    b.blockpart[((size_t)parity * nblk + blk) * kBP + 10 + q] = (double)t;
  }
}

template <int M, bool STAGE>
__global__ __launch_bounds__(512) void k_md_tile(ModelTable mt, Geom g, MdBuf b, PairK<float> pk,
                                                 int parity, CloseArgs ca, TileLayout lay) {
  if (ca.mode) {
    // block 0 closes the previous step while the other blocks run the tiles
    if (blockIdx.x == 0) {
      __shared__ double s_close[4][kRed];
      close_body<false>(mt, b, g.n, ca.dt, ca.nblk_pre, ca.bpart, ca.nblk_main, g.ncell,
                        ca.nose_active, ca.mode, parity, box_volume(g.L), s_close);
      return;
    }
    tile_body<M, STAGE>(g, b, pk, parity, lay, blockIdx.x - 1, gridDim.x - 1);
  } else {
    tile_body<M, STAGE>(g, b, pk, parity, lay, blockIdx.x, gridDim.x);
  }
}

// force (lab) and body torque of every molecule from its partials, in the
// order the corrector sums them (g6_eval_forces)
__global__ __launch_bounds__(256) void k_part_reduce(MdBuf b, int n, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double f[6] = {0, 0, 0, 0, 0, 0};
  const float *p = b.part + (size_t)i * b.nds * 6;
  for (int q = 0; q < b.nds; q++)
#pragma unroll
    for (int c = 0; c < 6; c++) f[c] += (double)p[q * 6 + c];
  double t[9];
#pragma unroll
  for (int a = 0; a < 9; a++) t[a] = b.tmat[a * n + i];
  out[i] = f[0];
  out[n + i] = f[1];
  out[2 * n + i] = f[2];
  out[3 * n + i] = t[0] * f[3] + t[3] * f[4] + t[6] * f[5];
  out[4 * n + i] = t[1] * f[3] + t[4] * f[4] + t[7] * f[5];
  out[5 * n + i] = t[2] * f[3] + t[5] * f[4] + t[8] * f[5];
}

TileLayout tile_layout(int nq, int nbmax, int cap, int waves, bool stage) {
  auto up = [](size_t x) { return (int)((x + 15) / 16 * 16); };
  TileLayout t{};
  t.nbmax = nbmax;
  int o = 0;
  t.cx = o; o += up(8 * (size_t)nbmax);
  t.cy = o; o += up(8 * (size_t)nbmax);
  t.cz = o; o += up(8 * (size_t)nbmax);
  t.id = o; o += up(4 * (size_t)nbmax);
  t.order = o; o += up(4 * (size_t)cap);
  t.site = o;
  if (stage) o += up(12 * (size_t)nq * nbmax);
  // accumulators, rings and row buffers last and contiguous: at the end
  // they take the 10 x 64 x waves lane partials of the energy and virial
  const size_t tail = (size_t)waves * (16 * kRing + 4 * kRowBuf);
  size_t accb = 48 * (size_t)nbmax;
  if (accb + tail < 5120 * (size_t)waves) accb = 5120 * (size_t)waves - tail;
  t.acc = o; o += up(accb);
  t.ring = o; o += waves * 16 * kRing;
  t.rowbuf = o; o += waves * 4 * kRowBuf;
  t.bytes = o;
  t.stage = stage ? 1 : 0;
  return t;
}

template <int M>
static void tile_t(System &s, int parity, const CloseArgs &ca) {
  const TileLayout &lay = s.tile_lay;
  const unsigned grid = s.nblk_md + (ca.mode ? 1 : 0);
  const unsigned threads = 64 * s.tile_waves;
  if (lay.stage)
    k_md_tile<M, true><<<grid, threads, lay.bytes, s.stream>>>(s.mt, s.geo_md, s.md, s.pkf,
                                                               parity, ca, lay);
  else
    k_md_tile<M, false><<<grid, threads, lay.bytes, s.stream>>>(s.mt, s.geo_md, s.md, s.pkf,
                                                                parity, ca, lay);
}

void md_launch_tile(System &s, int parity, const CloseArgs &ca) {
  if (s.kmodel == G6_TIP4P) tile_t<G6_TIP4P>(s, parity, ca);
  else if (s.kmodel == G6_TIP5P) tile_t<G6_TIP5P>(s, parity, ca);
  else if (s.kmodel == G6_SPCE) tile_t<G6_SPCE>(s, parity, ca);
  else if (s.kmodel == G6_NVDE) tile_t<G6_NVDE>(s, parity, ca);
  else if (s.kmodel == G6_OPLSMEOH) tile_t<G6_OPLSMEOH>(s, parity, ca);
  else tile_t<G6_ST2>(s, parity, ca);
}

void md_launch_part_reduce(System &s, double *out) {
  k_part_reduce<<<(s.n + 255) / 256, 256, 0, s.stream>>>(s.md, s.n, out);
}

// the kernels' dynamic LDS above the 64 KiB default
int tile_set_lds_limit(int bytes) {
  const void *fns[] = {
      (const void *)k_md_tile<G6_TIP4P, true>, (const void *)k_md_tile<G6_TIP4P, false>,
      (const void *)k_md_tile<G6_TIP5P, true>, (const void *)k_md_tile<G6_TIP5P, false>,
      (const void *)k_md_tile<G6_SPCE, true>,  (const void *)k_md_tile<G6_SPCE, false>,
      (const void *)k_md_tile<G6_NVDE, true>,  (const void *)k_md_tile<G6_NVDE, false>,
      (const void *)k_md_tile<G6_OPLSMEOH, true>, (const void *)k_md_tile<G6_OPLSMEOH, false>,
      (const void *)k_md_tile<G6_ST2, true>,   (const void *)k_md_tile<G6_ST2, false>};
  for (const void *f : fns)
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
      return G6_EHIP;
  return 0;
}

}  // namespace g6
