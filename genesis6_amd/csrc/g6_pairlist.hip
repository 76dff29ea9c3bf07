// The reference's cell grid and pair list on the GPU (the g6_pairlist entry
// point), reproducing the list element for element: the same pairs in the
// same order with the same FP64 image shifts.
//
//   Address_Assign      GridModule.F90:503-552   k_bin_only
//   residents in molecule order                  k_scan_cells, k_scatter,
//                                                k_sort_cells
//   Grid_NeighborList   GridModule.F90:557-752   k_rows (count / write)
//   Interaction_Compress2 InteractionModule.F90:398-605 (FP64 test dd < rr,
//                       fused into k_rows; shifts packed as integers)
//
// FP64 expressions are evaluated exactly as the reference writes them, so
// FMA contraction is off for this file.
#include <vector>

#include "g6_internal.h"

#pragma clang fp contract(off)

namespace g6 {

constexpr int kWave = 64;

// Orders LDS traffic among the lanes of one wave (no block barrier: the
// waves of a block work on independent rows and may exit early).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exclusive scan of in[0..m) into out[0..m] by one block of 1024 threads.
// copy (optional) receives out as well; total > cap raises flags[0].
__global__ __launch_bounds__(1024) void k_exscan(const int *in, int *out,
                                                 int *copy, int m,
                                                 const int *gate, int cap,
                                                 int *flags) {
  if (gate && !gate[0]) return;
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int chunk = (m + 1023) / 1024;
  const int b = t * chunk, e = min(m, b + chunk);
  int s = 0;
  for (int k = b; k < e; k++) s += in[k];
  part[t] = s;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    int v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int k = b; k < e; k++) {
    out[k] = run;
    if (copy) copy[k] = run;
    run += in[k];
  }
  if (t == 1023) {
    out[m] = part[1023];
    if (copy) copy[m] = part[1023];
    if (cap > 0 && part[1023] > cap) flags[0] = 1;
  }
}

// Unordered placement of each molecule into its cell segment; the order is
// restored by k_sort_cells.  Also records the COMs the list is built from.
__global__ void k_scatter(const int *cell, int *fill, int *tmp, const double *com,
                          double *com_ref, int n, const int *gate) {
  if (!gate[0]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int slot = atomicAdd(&fill[cell[i]], 1);
  tmp[slot] = i;
  com_ref[i] = com[i];
  com_ref[n + i] = com[n + i];
  com_ref[2 * n + i] = com[2 * n + i];
}

// One wave per cell: residents in ascending molecule index (the order in
// which Address_Assign appends them, GridModule.F90:538-548).
__global__ __launch_bounds__(256) void k_sort_cells(const int *start,
                                                    const int *tmp, int *sorted,
                                                    int ncell, const int *gate) {
  if (!gate[0]) return;
  const int c = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  if (c >= ncell) return;
  const int b = start[c], m = start[c + 1] - b;
  for (int t = lane; t < m; t += kWave) {
    const int v = tmp[b + t];
    int rank = 0;
    for (int u = 0; u < m; u++) rank += tmp[b + u] < v;
    sorted[b + rank] = v;
  }
}

__device__ __host__ inline int nbr_cell(const Geom &g, int c, int packed) {
  const int nx = g.nd[0], ny = g.nd[1], nz = g.nd[2];
  const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
  const int ox = packed & 0xff, oy = (packed >> 8) & 0xff, oz = (packed >> 16) & 0xff;
  return (((cz + oz) % nz) * ny + (cy + oy) % ny) * nx + (cx + ox) % nx;
}

// Shift code: three 10-bit fields holding k + 512 for ox = k*L, k in
// [-511, 511]; the field value 0 stands for -0.0 (dnint of a small negative
// difference), which the reference's FP64 shifts carry.
__device__ inline int shift_field(double k) {
  return (k == 0.0 && signbit(k)) ? 0 : (int)k + 512;
}
__device__ inline int shift_code(double kx, double ky, double kz) {
  return shift_field(kx) | (shift_field(ky) << 10) | (shift_field(kz) << 20);
}
__device__ __host__ inline double shift_value(int field, double L) {
  return field == 0 ? -0.0 : (double)(field - 512) * L;
}

// One wave per row r (r-th molecule in cell order).  mode 0 counts the
// compressed pairs of both parts of the row; mode 1 writes them at the
// scanned offsets, keeping the candidate order via ballot + prefix count.
template <int MODE>
__global__ __launch_bounds__(256) void k_rows(Geom g, const double *com,
                                              const int *cell, const int *start,
                                              const int *sorted, int *cnt,
                                              const int *off, int *pi, int *pj,
                                              int *scode, int pcap,
                                              const int *gate, int *flags) {
  if (!gate[0]) return;
  __shared__ int s_beg[4][32], s_pre[4][33];
  const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int r = blockIdx.x * 4 + wid;
  if (r >= g.n) return;
  const int n = g.n;
  const int i = sorted[r];
  const int c = cell[i];
  const double xi = com[i], yi = com[n + i], zi = com[2 * n + i];
  for (int part = 0; part < 2; part++) {
    const int nst = part == 0 ? g.nst : g.nst0;
    int len = 0;
    if (lane < nst) {
      const int c2 = nbr_cell(g, c, part == 0 ? g.st[lane] : g.st0[lane]);
      s_beg[wid][lane] = start[c2];
      len = start[c2 + 1] - start[c2];
    }
    // inclusive wave scan of len over the stencil entries
    int inc = len;
    for (int d = 1; d < 32; d <<= 1) {
      const int v = __shfl_up(inc, d, kWave);
      if (lane >= d) inc += v;
    }
    if (lane < nst) s_pre[wid][lane + 1] = inc;
    if (lane == 0) s_pre[wid][0] = 0;
    wave_sync();
    const int total = s_pre[wid][nst];
    int base = MODE == 1 ? off[part * n + r] : 0;
    int count = 0;
    for (int t0 = 0; t0 < total; t0 += kWave) {
      const int t = t0 + lane;
      bool pass = false;
      int j = 0, code = 0;
      if (t < total) {
        int s = 0;
        while (s + 1 < nst && s_pre[wid][s + 1] <= t) s++;
        j = sorted[s_beg[wid][s] + (t - s_pre[wid][s])];
        if (part == 0 || j > i) {
          double d0 = xi - com[j], d1 = yi - com[n + j], d2 = zi - com[2 * n + j];
          const double k0 = round(d0 * g.invL[0]), k1 = round(d1 * g.invL[1]),
                       k2 = round(d2 * g.invL[2]);
          d0 = d0 - k0 * g.L[0];
          d1 = d1 - k1 * g.L[1];
          d2 = d2 - k2 * g.L[2];
          const double dd = d0 * d0 + d1 * d1 + d2 * d2;
          pass = dd < g.rr;
          if (MODE == 1 && pass) {
            if (fabs(k0) > 511.0 || fabs(k1) > 511.0 || fabs(k2) > 511.0) flags[0] = 2;
            code = shift_code(k0, k1, k2);
          }
        }
      }
      const unsigned long long mask = __ballot(pass);
      if (MODE == 1) {
        const int pos = base + __popcll(mask & ((1ull << lane) - 1ull));
        if (pass && pos < pcap) {
          pi[pos] = i;
          pj[pos] = j;
          scode[pos] = code;
        }
        base += __popcll(mask);
      } else {
        count += __popcll(mask);
      }
    }
    if (MODE == 0 && lane == 0) cnt[part * n + r] = count;
    wave_sync();
  }
}

// Host-array entry: the reference's list for given COMs, written to caller
// arrays as 0-based indices and FP64 shifts.
__global__ void k_bin_only(Geom g, const double *com, int *cell, int *count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int n = g.n;
  int jc[3];
  for (int a = 0; a < 3; a++) {
    double d = com[a * n + i];
    const double o = round(d * g.invL[a]) * g.L[a];
    d = d - o;
    int j = (int)round(d * g.cellinv[a]);
    if (j < 0) j += g.nd[a];
    // the reference indexes out of range here only for ndiv = 1
    j = ((j % g.nd[a]) + g.nd[a]) % g.nd[a];
    jc[a] = j;
  }
  const int c = (jc[2] * g.nd[1] + jc[1]) * g.nd[0] + jc[0];
  cell[i] = c;
  atomicAdd(&count[c], 1);
}

int cells_once(System &s, const double *com_h, int32_t *cell_h) {
  const Geom &g = s.geo;
  char *base = nullptr;
  const int rs = seam_reserve(s, seam_piece(sizeof(double) * 3 * g.n) +
                                     seam_piece(sizeof(int) * (g.n + g.ncell + 1)),
                              &base);
  if (rs != 0) return rs;
  Carve cv{base};
  double *com = cv.take<double>(3 * (size_t)g.n);
  int *cell = cv.take<int>((size_t)g.n + g.ncell + 1);
  int *count = cell + g.n;
  G6_CHECK(hipMemcpyAsync(com, com_h, sizeof(double) * 3 * g.n, hipMemcpyHostToDevice, s.stream));
  G6_CHECK(hipMemsetAsync(count, 0, sizeof(int) * (g.ncell + 1), s.stream));
  k_bin_only<<<(g.n + 255) / 256, 256, 0, s.stream>>>(g, com, cell, count);
  G6_CHECK(hipGetLastError());
  G6_CHECK(hipMemcpyAsync(cell_h, cell, sizeof(int) * g.n, hipMemcpyDeviceToHost, s.stream));
  G6_CHECK(hipStreamSynchronize(s.stream));
  return 0;
}

__global__ void k_unpack_shifts(const int *code, int np, double Lx, double Ly, double Lz,
                                double *ox, double *oy, double *oz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int v = code[k];
  ox[k] = shift_value(v & 1023, Lx);
  oy[k] = shift_value((v >> 10) & 1023, Ly);
  oz[k] = shift_value((v >> 20) & 1023, Lz);
}

int pairlist_once(System &s, const double *com_h, int64_t cap, int64_t *npair0,
                  int64_t *npair, int32_t *pi_h, int32_t *pj_h, double *ox_h,
                  double *oy_h, double *oz_h) {
  const Geom &g = s.geo;
  const int n = g.n;
  // a private list build on the caller's COMs, into the system's buffers
  char *base = nullptr;
  const int rs = seam_reserve(s, 2 * seam_piece(sizeof(double) * 3 * n), &base);
  if (rs != 0) return rs;
  Carve cv{base};
  double *com = cv.take<double>(3 * (size_t)n), *ref = cv.take<double>(3 * (size_t)n);
  G6_CHECK(hipMemcpyAsync(com, com_h, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                          s.stream));
  G6_CHECK(hipMemsetAsync(s.cell_count, 0, sizeof(int) * (g.ncell + 1), s.stream));
  int one = 1;
  G6_CHECK(hipMemcpyAsync(s.lflags + 1, &one, sizeof(int), hipMemcpyHostToDevice, s.stream));
  G6_CHECK(hipMemsetAsync(s.lflags, 0, sizeof(int), s.stream));
  k_bin_only<<<(n + 255) / 256, 256, 0, s.stream>>>(g, com, s.cell, s.cell_count);
  const int *gate = s.lflags + 1;
  Geom g2 = g;
  k_exscan<<<1, 1024, 0, s.stream>>>(s.cell_count, s.cell_start, s.cell_fill, g.ncell,
                                     gate, 0, s.lflags);
  k_scatter<<<(n + 255) / 256, 256, 0, s.stream>>>(s.cell, s.cell_fill, s.sorted_tmp, com,
                                                    ref, n, gate);
  k_sort_cells<<<(g.ncell + 3) / 4, 256, 0, s.stream>>>(s.cell_start, s.sorted_tmp,
                                                         s.sorted, g.ncell, gate);
  const int nb = (n + 3) / 4;
  k_rows<0><<<nb, 256, 0, s.stream>>>(g2, com, s.cell, s.cell_start, s.sorted, s.cnt,
                                      s.off, s.pi, s.pj, s.scode,
                                      s.pcap, gate, s.lflags);
  k_exscan<<<1, 1024, 0, s.stream>>>(s.cnt, s.off, nullptr, 2 * n, gate, s.pcap, s.lflags);
  k_rows<1><<<nb, 256, 0, s.stream>>>(g2, com, s.cell, s.cell_start, s.sorted, s.cnt,
                                      s.off, s.pi, s.pj, s.scode,
                                      s.pcap, gate, s.lflags);
  int total = 0, flag = 0;
  G6_CHECK(hipMemcpyAsync(&total, s.off + 2 * n, sizeof(int), hipMemcpyDeviceToHost,
                          s.stream));
  G6_CHECK(hipMemcpyAsync(&flag, s.lflags, sizeof(int), hipMemcpyDeviceToHost, s.stream));
  G6_CHECK(hipStreamSynchronize(s.stream));
  // candidates: sum over rows of the stencil residents (counted on host from
  // the cell occupancy; cheap and exact)
  {
    std::vector<int> st(g.ncell + 1), cl(n), so(n);
    G6_CHECK(hipMemcpy(st.data(), s.cell_start, sizeof(int) * (g.ncell + 1),
                       hipMemcpyDeviceToHost));
    G6_CHECK(hipMemcpy(cl.data(), s.cell, sizeof(int) * n, hipMemcpyDeviceToHost));
    G6_CHECK(hipMemcpy(so.data(), s.sorted, sizeof(int) * n, hipMemcpyDeviceToHost));
    int64_t c0 = 0;
    auto nb_host = [&](int c, int packed) { return nbr_cell(g, c, packed); };
    for (int r = 0; r < n; r++) {
      const int i = so[r], c = cl[i];
      for (int t = 0; t < g.nst; t++) {
        const int c2 = nb_host(c, g.st[t]);
        c0 += st[c2 + 1] - st[c2];
      }
      for (int t = 0; t < g.nst0; t++) {
        const int c2 = nb_host(c, g.st0[t]);
        for (int q = st[c2]; q < st[c2 + 1]; q++) c0 += so[q] > i;
      }
    }
    *npair0 = c0;
  }
  *npair = total;
  int rc = 0;
  if (flag == 2) rc = G6_EARG;
  if (total > cap || total > s.pcap) rc = G6_ECAPACITY;
  if (rc == 0 && total > 0) {
    // com and ref are dead here (the stream has been synchronized)
    char *ob = nullptr;
    const int ro = seam_reserve(s, seam_piece(sizeof(double) * 3 * (size_t)total), &ob);
    if (ro != 0) return ro;
    double *o = reinterpret_cast<double *>(ob);
    k_unpack_shifts<<<(total + 255) / 256, 256, 0, s.stream>>>(
        s.scode, total, g.L[0], g.L[1], g.L[2], o, o + total, o + 2 * (size_t)total);
    G6_CHECK(hipMemcpyAsync(pi_h, s.pi, sizeof(int) * total, hipMemcpyDeviceToHost, s.stream));
    G6_CHECK(hipMemcpyAsync(pj_h, s.pj, sizeof(int) * total, hipMemcpyDeviceToHost, s.stream));
    G6_CHECK(hipMemcpyAsync(ox_h, o, sizeof(double) * total, hipMemcpyDeviceToHost, s.stream));
    G6_CHECK(hipMemcpyAsync(oy_h, o + total, sizeof(double) * total, hipMemcpyDeviceToHost,
                            s.stream));
    G6_CHECK(hipMemcpyAsync(oz_h, o + 2 * (size_t)total, sizeof(double) * total,
                            hipMemcpyDeviceToHost, s.stream));
    G6_CHECK(hipStreamSynchronize(s.stream));
  }
  // the MD list in the system buffers is no longer valid
  G6_CHECK(hipMemcpy(s.lflags + 1, &one, sizeof(int), hipMemcpyHostToDevice));
  G6_CHECK(hipMemset(s.cell_count, 0, sizeof(int) * (g.ncell + 1)));
  return rc;
}

}  // namespace g6
