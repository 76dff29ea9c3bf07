// Internal definitions shared by the g6gpu translation units.
//
// Device data layout (one system per handle, all resident in HBM):
//   rigid state, FP64, component-major:  com[3][n], dx[5][3][n], w[5][3][n],
//     quat[5][4][n]  (the Nordsieck arrays of RigidModule.F90:14-33 sRigid1)
//   tmat[9][n]        rotation matrices (rigid_setmatrix, RigidModule.F90:62-83)
//   intra[ns][3][n]   site offsets from the COM, FP64; intraf the FP32 copy
//   cell[n], cell_count/start/fill[ncell+1], sorted[n]: cell residents in
//     molecule order (Address_Assign, GridModule.F90:503-552)
//   row counts cnt[2n] and offsets off[2n+1]: row r = r-th molecule in cell
//     order; rows 0..n-1 hold the half-shell part of the candidate list,
//     rows n..2n-1 the self-mirror part (Grid_NeighborList,
//     GridModule.F90:557-752); pairs in [off[r], off[r+1])
//   pj[pcap], pi[pcap], scode[pcap]: compressed pair list; scode packs the
//     image shift ox = kx*L etc. as 10-bit fields (kx+512) | (ky+512)<<10 |
//     (kz+512)<<20 (COMs are unwrapped, so |k| can exceed 1); field value 0
//     encodes the -0.0 shift that dnint gives for small negative differences
//   facc[12][n]       FP64 accumulators: [0..5] force and torque (lab) that
//     molecule i receives as the j of other rows (atomics), [6..11] its own
//     row's share (plain stores, one writer)
//   rowpart[11][n]    per-row energy, virial, interacting-pair count (FP64)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/g6gpu.h"

namespace g6 {

constexpr int kMaxSite = 6;
constexpr int kMaxCoul = 16;

// Interaction table of one rigid model, computed on the host with the
// reference's coefficient arithmetic and passed to kernels by value.
struct ModelTable {
  int model;
  int nsite;     // including the CoM site (last)
  int ncoul;     // Coulomb site pairs
  int ci[kMaxCoul], cj[kMaxCoul];
  double qq[kMaxCoul];  // prefactor, internal units
  int lj_site;          // LJ site (O = 0), O-O only for all three models
  double lj_a, lj_b;    // e = a/r^12 + b/r^6
  double bx[kMaxSite], by[kMaxSite], bz[kMaxSite];  // body-frame sites
  double Ixx, Iyy, Izz, mass;
  double rl, ru;  // ST2 short-range switch on the COM distance
};

// Grid geometry and cutoff constants of one system.
struct Geom {
  int n, ndiv, ncell;
  int nst, nst0;           // half-shell and self-mirror stencil sizes
  int st[27], st0[27];     // stencil offsets packed ox | oy<<8 | oz<<16
  double L, invL, cellinv; // box edge, 1/L, 1/(L/ndiv) as the reference
  double rr;               // pair-list radius squared ((rc+skin)^2)
  double rc2;              // rc^2 (interaction filter)
  double rin, rout, c1, c2; // switch: S between rin and rout
};

struct System {
  int model = 0, n = 0, ns = 0, prec = 0;
  double L = 0, rc = 0, taper = 0, req = 0, dt = 0, skin = 0;
  ModelTable mt{};
  Geom geo{};
  int pcap = 0;
  bool nose_active = false;
  double nose_h[9] = {0};
  // device buffers
  double *com = nullptr, *dx = nullptr, *w = nullptr, *quat = nullptr;
  double *tmat = nullptr, *intra = nullptr;
  float *intraf = nullptr;
  double *com_ref = nullptr;  // COMs at the last list build (skin > 0)
  int *cell = nullptr, *cell_count = nullptr, *cell_start = nullptr,
      *cell_fill = nullptr, *sorted_tmp = nullptr, *sorted = nullptr;
  int *cnt = nullptr, *off = nullptr;
  int *pi = nullptr, *pj = nullptr;
  int *scode = nullptr;
  double *facc = nullptr, *rowpart = nullptr, *blockpart = nullptr;
  double *nose = nullptr, *thermo = nullptr;
  int *flags = nullptr;  // [0] overflow, [1] rebuild request, [2] builds
  int nblk_mol = 0;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0;
  hipEvent_t ev[8];
  double t_force = 0, t_list = 0, t_integ = 0, t_step = 0;
  int steps_done = 0;
};

// model tables (g6_api.hip)
bool make_model_table(int model, int prec, ModelTable *mt);

// launchers (stream-ordered, graph-capturable: no memset, no host sync)
void launch_predict_sites_bin(System &s, bool bin);
void launch_pairlist(System &s);
void launch_force_rows(System &s);
void launch_correct(System &s);
void launch_thermo(System &s);
void launch_collect_acc(System &s, double *out);
// host-array entry points of the module seam
int pairlist_once(System &s, const double *com_h, int64_t cap, int64_t *npair0,
                  int64_t *npair, int32_t *pi, int32_t *pj, double *ox,
                  double *oy, double *oz);
int cells_once(System &s, const double *com_h, int32_t *cell);
int cutoff_once(System &s, int64_t np, const int32_t *pi, const int32_t *pj,
                const double *com, const double *ox, const double *oy,
                const double *oz, double *er, double *fr);
int sites_once(System &s, const double *com_h, const double *q_h, double *x,
               double *y, double *z);
int collect_once(System &s, const double *q_h, const double *fx, const double *fy,
                 const double *fz, double *force, double *torque);
int force_pairs(System &s, int64_t np, const int32_t *pi, const int32_t *pj,
                const double *ox, const double *oy, const double *oz,
                const double *er, const double *fr, const double *x,
                const double *y, const double *z, double *fx, double *fy,
                double *fz, double *ep, double *vrmat, bool energy_only);

}  // namespace g6

#define G6_CHECK(x)                       \
  do {                                    \
    hipError_t e_ = (x);                  \
    if (e_ != hipSuccess) return G6_EHIP; \
  } while (0)
