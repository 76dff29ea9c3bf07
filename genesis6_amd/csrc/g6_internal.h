// Internal definitions shared by the g6gpu translation units.
//
// Device data layout (one system per handle, all resident in HBM).
//
// MD path (g6_md.hip), structure MdBuf:
//   rigid state, FP64, component-major:  com[3][n], dx[5][3][n], w[5][3][n],
//     quat[5][4][n]  (the Nordsieck arrays of RigidModule.F90:14-33 sRigid1)
//   tmat[9][n]        rotation matrices (rigid_setmatrix, RigidModule.F90:62-83)
//   rec[n][rs]        per-molecule record of the predicted COM and site offsets
//   cellres[2][ncell][cell_cap]: fixed-capacity cell lists of the MD grid
//     (cell edge >= (rc + skin)/2), one per step parity; slot 0 holds the
//     count so that no two cells share a counter's cache line
//   nbr_j/nbr_code[n][maxnb], nbr_cnt[n]: full (both-direction) Verlet rows
//     within rc + skin; nbr_code packs the image shift ox = kx*L etc. as
//     10-bit fields (kx+512) | (ky+512)<<10 | (kz+512)<<20 (COMs are
//     unwrapped, so |k| can exceed 1).  The -0.0 shift that dnint gives for
//     small negative differences is stored as 0: x - (-0.0) and x - 0.0 are
//     equal for every x but a zero x, whose sign no force term can see
//   blockpart[2][nblk_md][16], blockek[pre waves][10]: per-block energy/virial
//     and per-wave kinetic partials, reduced in a fixed order by the step
//     close (one block)
// Reference-list entry point (g6_pairlist.hip): cell residents in molecule
//   order and the half-shell candidate rows of Grid_NeighborList
//   (GridModule.F90:557-752) with counts cnt[2n], offsets off[2n+1], and the
//   compressed list pi/pj/scode[pcap] in the reference's order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/g6gpu.h"

namespace g6 {

constexpr int kMaxSite = 7;
constexpr int kMaxCoul = 25;
constexpr int kMaxLJ = 8;  // LJ site pairs besides O-O

// Interaction table of one rigid model, computed on the host with the
// reference's coefficient arithmetic and passed to kernels by value.
struct ModelTable {
  int model;
  int nsite;     // including the CoM site (last)
  int ncoul;     // Coulomb site pairs
  int ci[kMaxCoul], cj[kMaxCoul];
  double qq[kMaxCoul];  // prefactor, internal units
  int lj_site;          // the O-O LJ term (site 0 with site 0)
  double lj_a, lj_b;    // e = a/r^12 + b/r^6
  // further LJ site pairs (NvdE: O-H, H-O, H-H), coefficients as lj_a, lj_b
  int nlx;
  double lx_a[kMaxLJ], lx_b[kMaxLJ];
  double bx[kMaxSite], by[kMaxSite], bz[kMaxSite];  // body-frame sites
  double Ixx, Iyy, Izz, mass;
  double inv_ixx, inv_iyy, inv_izz, inv_mass;  // 1/I and 1/m, rounded on the host
  double rl, ru;  // ST2 short-range switch on the COM distance
  // StdInteraction site parameters (si%param: charge, LJ eps, LJ sigma) of
  // TIP4P_SetInteraction / TIP5P_SetInteraction, for the energy-only path
  double std_q[kMaxSite], std_eps[kMaxSite], std_sig[kMaxSite];
};

// StdPairPotential site terms (StdInteraction.F90:633-762), site a of i
// outer and b of j inner; kind 1 LJ+Coulomb, 2 Coulomb, 3 LJ.
struct StdTerm {
  int a, b, kind;
  double qq, aa, bb;
};
constexpr int kMaxStd = kMaxSite * kMaxSite;
struct StdTable {
  int nt;
  StdTerm t[kMaxStd];
};

// Pair constants of the MD force kernel in its arithmetic type R, rounded
// once on the host (kernel arguments, so they sit in scalar registers).
template <typename R>
struct PairK {
  R qq[kMaxCoul];
  R lja, ljb, lja12, ljb6;  // e = a/r^12 + b/r^6; force factors 12a, 6b
  R xa[kMaxLJ], xb[kMaxLJ], xa12[kMaxLJ], xb6[kMaxLJ];  // the further LJ terms
  R rin, rout, rin2, rout2, c1, c2;  // switch S between rin and rout
  R rl, ru, rl2, ru2, w3inv;         // ST2 switch on the COM distance
};

constexpr int kBoxLen = 17;  // doubles of the barostat's device box

// Grid geometry and cutoff constants of one system.  The box is
// orthorhombic: edges L[a], 1/L[a] and 1/(L[a]/nd[a]) per axis, as the
// reference's sBox and sGrid hold them (BoxModule.F90, GridModule.F90:473-501).
struct Geom {
  int n, ncell;
  int nd[3];               // cells per axis
  int nst, nst0;           // half-shell and self-mirror stencil sizes
  int st[27], st0[27];     // stencil offsets packed ox | oy<<8 | oz<<16
  double L[3], invL[3], cellinv[3];
  double rr;               // pair-list radius squared ((rc+skin)^2)
  double rc2;              // rc^2 (interaction filter)
  double rin, rout, c1, c2; // switch: S between rin and rout
};

// Per-molecule record strides (in doubles) read by k_md_main: the predicted
// COM in FP64 at [0..2], then the site offsets of the sites before the CoM
// site, as floats from [4] (FP32 mode, 128 B) or doubles from [4] (FP64).
constexpr int kRecF = 16, kRecD = 24;

// Device pointers and sizes of the MD path, passed to kernels by value.
struct MdBuf {
  double *com, *dx, *w, *quat, *tmat, *com_ref;
  double *rec;  // [n][rs] records
  int rs;
  int *cell;     // [n] MD-grid cell of each molecule
  int *cellres;  // [2][ncell][cell_cap]: slot 0 the count, residents from slot 1
  int cell_cap;
  int *nbr_j, *nbr_code, *nbr_cnt;
  int maxnb;
  double *frc;        // [6][n] force and torque (lab) of the last k_md_main
  double *blockpart;  // [2][nblk][16] k_md_main block partials, by parity
  double *blockek;    // [k_md_pre waves][kEkLen] kinetic partials
  double *nose, *thermo;
  int *flags;  // [0] error code, [2 + parity] rebuild rows this step
  double half_skin2;
  // Andersen constant-pressure MD (both nullptr for a fixed box).  box
  // (kBoxLen doubles): [0..2] the current edges, [3..5] their inverses,
  // [6..8] 1/(L/nd) of the MD grid per axis, [9..11] the edges at the last
  // row build, [12..14] nd per axis, [15] the smallest cell edge the grid
  // allows, [16] the barostat mode (1 orthorhombic: every edge follows the
  // volume; 2 orz: the z edge alone, AndersenModule.F90:65-82).
  // anders: pex (atm), mass (internal units), v0..v5 (AndersenModule.F90:5-8)
  double *box, *anders;
  // Cell tiles (g6_tile.hip), FP32 fixed-box systems whose grid has cells of
  // edge >= rc + skin; part == nullptr selects the rows of both directions.
  //   part[n][nds][6]    force and lab torque partials: each tile writes one
  //                      slot for every molecule of its neighbourhood
  //   nb_id[ncell][nbmax] molecule ids of a tile's neighbourhood (its own cell,
  //                      then the 13 forward cells, each sorted by id), and
  //   nb_base[ncell][16] the first slot of each of the 14 cells ([14] total),
  //                      both frozen at the last row build
  //   trow[n][maxrow]    half rows: slot | image code << 11; trow_cnt[n]
  //   tcnt[ncell][cell_cap] the row length of each tile member
  float *part;
  int nds, slices, nbmax, maxrow;
  int *nb_id, *nb_base, *trow_cnt, *tcnt;
  unsigned *trow;
};

// Descriptor of one system of a ladder, in device memory: the ladder
// kernels (k_md_*_lad) read it per block instead of taking it by value.
struct MdRep {
  Geom g;  // the MD grid of the system
  MdBuf b;
  double dt;
  int nose_active, nblk_main;
};

// Monte Carlo state (g6_mc.hip): the mc2 single-molecule trial loop on a
// configuration of its own, FP64 throughout.
//   com[3][n], quat[4][n] (order-1 quaternions), sx/sy/sz[n*ns] site
//   positions, pinfo[n][n] PairInfo pair energies (symmetric), delta[n] and
//   rtr[n] step sizes, u[ucap] the Random_GetNext stream of this call and
//   z[ucap] the Box-Muller factor sqrt(-2 ln u[k]) sin(2 pi u[k+1]) of each
//   position, pos/nacc the stream position and accept count, epsum.
// Umbrellas of the MC trial loop: potentials coeffi (op - center)^2 on the
// triangle histogram's divergence (kinds 1 and 4, log for 4) or the energy
// (kind 3).  The triangle tables hold, per molecule, the molecules within
// 5 A at the start of the call (kTriNear entries at most): label and the
// vector to it with its squared length (dx, dy, dz, dd).
constexpr int kUmbMax = 4;
constexpr int kTriNear = 64;
struct McUmb {
  int numb = 0;
  int kind[kUmbMax] = {0, 0, 0, 0};
  double center[kUmbMax] = {0, 0, 0, 0}, coeffi[kUmbMax] = {0, 0, 0, 0};
  double refhist[21] = {};
  int tri = 0;  // any triangle kind
  int q6 = 0;   // any Q6LOCAL_ (kind 2)
};

// The Q6LOCAL_ umbrella's state (the reference's sq6system for one loop):
// rows within 6 A by label (dx, dy, dz, dd per entry; the entries within 3
// and 5 A are nei3 and nei5), the ordered sets neis(i) and neib(jj, i) of
// every bond, qlm of every bond, qlmsum and nbond of every molecule, each
// molecule's bond count, q6 of the trial, and the order parameter prepared
// at the loop's start.
constexpr int kQ6Near = 64;  // entries of a 6 A row
constexpr int kQ6Set = 128;  // elements of a set
constexpr int kQ6Rec = 32;   // doubles of a qlm record (26 used)
struct Q6Buf {
  int *nn6 = nullptr, *lab6 = nullptr;
  double *v6 = nullptr;
  int *neis_n = nullptr, *neis = nullptr;
  int *neib_n = nullptr, *neib = nullptr;
  double *qlm = nullptr, *qsum = nullptr, *dq6 = nullptr, *op = nullptr;
  int *nbond = nullptr, *nb3 = nullptr;
};

struct McBuf {
  double *com = nullptr, *quat = nullptr, *sx = nullptr, *sy = nullptr, *sz = nullptr;
  double *pinfo = nullptr, *delta = nullptr, *rtr = nullptr;
  double *u = nullptr, *z = nullptr, *epsum = nullptr, *part = nullptr;
  int *pos = nullptr;
  long long *nacc = nullptr;
  int ucap = 0;
  // umbrella state: triangle tables [n][kTriNear], histogram, the order
  // parameters after the last call, the capacity flag of the tables
  int *tri_nn = nullptr, *tri_near = nullptr, *tri_hist = nullptr, *uflag = nullptr;
  double *tri_v = nullptr, *uop = nullptr;
  Q6Buf q6;
};

// Host side of the MC random stream: the reference's MT19937 (mtrnd.f) and
// the Random_GetNext values generated but not yet consumed by the device.
struct McHost {
  uint32_t mt[624];
  int mti = 625;
  std::vector<double> pend;
  void seed(uint32_t s);
  uint32_t next();
  double getnext();  // 1 - the 31-bit uniform of mtrndv (RandomModule.F90:123-130)
};

// LDS layout of the tile kernel (byte offsets into its dynamic LDS): per slot
// the FP64 COM components, six 64-bit accumulators, the molecule id and
// (stage != 0) the FP32 site offsets; then per wave a ring of pair entries and
// a buffer for the next member's row.
struct TileLayout {
  int nbmax, cx, cy, cz, acc, id, site, ring, rowbuf, order, bytes, stage;
  int xcd;  // workgroups of one XCD take a contiguous run of tiles
};
TileLayout tile_layout(int nq, int nbmax, int cap, int waves, bool stage);

struct System {
  int model = 0, n = 0, ns = 0, prec = 0;
  // fast: the reference's @FAST (custom kernels for TIP4P and ST2); kmodel:
  // the pair-kernel layout the MD and force paths instantiate (ST2 without
  // fast runs StdForce, whose site terms are those of the TIP5P layout)
  int fast = 1, kmodel = 0;
  // Lb: the box edges; L = Lb[0], the edge of the paths that take cubic
  // boxes only (tiles, ladders, Monte Carlo, NPT Monte Carlo)
  double Lb[3] = {0, 0, 0};
  double L = 0, rc = 0, taper = 0, req = 0, dt = 0, skin = 0;
  ModelTable mt{};
  Geom geo{};     // the reference grid (ndiv = int(L/voxa)), entry points
  Geom geo_md{};  // the MD grid (cell edge >= rc + skin)
  MdBuf md{};
  McBuf mc{};
  bool mc_ready = false;
  McHost mch;
  double mc_beta = 0;  // 1/T of the MC run
  McUmb mc_umb;
  StdTable std{};
  int pcap = 0;
  bool nose_active = false;
  double nose_h[9] = {0};
  bool andersen = false;  // md.box and md.anders are allocated
  int andersen_mode = 0;  // 1 orthorhombic, 2 orz
  bool tiles = false;     // the MD path runs the cell-tile force kernel
  int tile_exit = 0;      // flag code of the capacity error that left the tiles
  int last_flag = 0;      // flag code the last check of the step flags read
  // a tile system moved to the rows by a capacity error tries the tiles
  // again at the first g6_step call from step tile_retry_at on; the gap
  // doubles after every failed try
  int tile_retry_at = 0, tile_retry_gap = 0;
  // device scratch of the host-array entry points, grown on demand
  char *seam = nullptr;
  size_t seam_cap = 0;
  bool tile_ok = false;   // tile buffers allocated (the system qualifies)
  int tile_waves = 8;     // waves per tile workgroup
  int tile_ndiv = 0, tile_cap = 0;  // the tile grid and its cell capacity
  TileLayout tile_lay{};
  float *tile_part = nullptr;  // md.part while the tiles run; kept for freeing
  double *snap = nullptr;      // the state at the start of a g6_step call
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
  double *nose = nullptr, *thermo = nullptr;
  int *flags = nullptr;   // = md.flags
  int *lflags = nullptr;  // g6_pairlist: [0] error code, [1] build gate
  int nblk_md = 0;
  PairK<float> pkf{};
  PairK<double> pkd{};
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;  // graph_steps steps
  int graph_steps = 0;
  // graphs of shorter runs, by step count (g6_step of any count replays)
  hipGraphExec_t graph_k[4] = {nullptr, nullptr, nullptr, nullptr};
  int graph_k_steps[4] = {0, 0, 0, 0};
  int graph_k_next = 0;
  int graph_gen = 0;  // drop_graphs calls so far (graphs of others that embed this system)
  hipEvent_t ev[8];
  double t_force = 0, t_list = 0, t_integ = 0, t_step = 0;
  int steps_done = 0;
};

// model tables (g6_api.hip)
bool make_model_table(int model, int prec, ModelTable *mt, bool fast = true);
// the pair-kernel layout of a model with or without the custom kernels
inline int kernel_model(int model, bool fast) { return model == G6_ST2 && !fast ? G6_TIP5P : model; }
void make_std_table(const ModelTable &mt, StdTable *t);
void make_std_table(const ModelTable &mi, const ModelTable &mj, StdTable *t);
// Rigid_VReset2 of each system in turn on one random stream (g6_api.hip)
int vreset_systems(System *const *sys, int nsys, double temp, uint32_t seed);
void mc_free(System &s);
void npt_forget(System *s);  // frees the NPT Monte Carlo state of s (g6_npt.hip)
System *system_of(g6_handle h);  // nullptr for an unknown handle
template <typename R>
PairK<R> make_pair_k(const ModelTable &mt, const Geom &g) {
  PairK<R> k{};
  for (int t = 0; t < kMaxCoul; t++) k.qq[t] = R(mt.qq[t]);
  k.lja = R(mt.lj_a);
  k.ljb = R(mt.lj_b);
  k.lja12 = R(mt.lj_a * 12.0);
  k.ljb6 = R(mt.lj_b * 6.0);
  for (int t = 0; t < kMaxLJ; t++) {
    k.xa[t] = R(mt.lx_a[t]);
    k.xb[t] = R(mt.lx_b[t]);
    k.xa12[t] = R(mt.lx_a[t] * 12.0);
    k.xb6[t] = R(mt.lx_b[t] * 6.0);
  }
  k.rin = R(g.rin);
  k.rout = R(g.rout);
  k.rin2 = R(g.rin * g.rin);
  k.rout2 = R(g.rout * g.rout);
  k.c1 = R(g.c1);
  k.c2 = R(g.c2);
  k.rl = R(mt.rl);
  k.ru = R(mt.ru);
  k.rl2 = k.rl * k.rl;
  k.ru2 = k.ru * k.ru;
  const R w = k.ru - k.rl;
  k.w3inv = w > R(0) ? R(1) / (w * w * w) : R(0);
  return k;
}

// MD launchers (stream-ordered, graph-capturable: no memset, no host sync)
// Arguments of the step-close work (k_md_close, or block 0 of k_md_main).
struct CloseArgs {
  int mode, nose_active, nblk_pre, nblk_main;
  const double *bpart;  // pair partials of the k_md_main being closed
  double dt;
};
// Bytes of the per-wave LDS region of k_md_main: the row sort buffers while
// a row is rebuilt, then the compacted interacting entries (index and the
// COM separation in the pair precision), then the lane partials of the
// energy and virial.
inline int main_region_bytes(int maxnb, int rsize) {
  auto up16 = [](int x) { return (x + 15) / 16 * 16; };
  const int sort = 8 * maxnb;
  const int comp = up16(4 * maxnb) + 3 * up16(rsize * maxnb);
  const int acc = 10 * 64 * 8;
  const int mx = sort > comp ? sort : comp;
  return up16(mx > acc ? mx : acc);
}
// doubles per block partial of k_md_main: the energy and the pair virial
// [0..9] (each a compensated block sum, rounded once), pair counts [10..11]
constexpr int kBP = 16;
// kinetic partials per k_md_pre wave: sum |dx(1)|^2, sum w_x^2, w_y^2, w_z^2,
// and the upper triangle xx xy xz yy yz zz of sum dx(1) dx(1)^T
constexpr int kEkLen = 10;
// reduction slots of the step close per wave: 12 pair quantities and kEkLen
// kinetic ones
constexpr int kRed = 12 + kEkLen;
// waves of k_md_pre (64 molecules per block of four waves); each writes
// one kinetic partial
inline int pre_waves(int n) { return (n + 63) / 64 * 4; }
void md_launch_pre(System &s, bool correct, bool predict, bool sites, int parity);
// flags[k] = v, stream-ordered
void md_launch_flag(System &s, int k, int v);
// close of the step whose k_md_main wrote the partials of src_parity
void md_launch_close(System &s, int mode, int parity, int src_parity);
// k_md_main of this parity; close_mode != 0 closes the previous step
// (partials of parity 1 - parity) in block 0
void md_launch_main(System &s, double *out, int parity, int close_mode);
// the cell-tile force kernel (g6_tile.hip) and the reduction of its partials
// into force and body torque (out[6][n])
void md_launch_tile(System &s, int parity, const CloseArgs &ca);
void md_launch_part_reduce(System &s, double *out);
int tile_set_lds_limit(int bytes);
// switch a tile system back to the rows of both directions (g6_api.hip)
int use_row_grid(System &s);
// destroys the captured step graphs (their launches embed the arguments)
void drop_graphs(System &s);
int graph_steps();  // steps per captured graph (G6_GRAPH_STEPS, default 50; 0: none)

// The host-array entry points stage their arrays in one device buffer per
// handle (grown on demand, freed by g6_finalize), so that no call allocates
// in the steady state and no error path leaks.  seam_reserve waits for the
// handle's stream before it replaces a smaller buffer.
int seam_reserve(System &s, size_t bytes, char **base);
inline size_t seam_piece(size_t bytes) { return ((bytes ? bytes : 1) + 255) / 256 * 256; }
struct Carve {
  char *p;
  template <class T>
  T *take(size_t count) {
    T *r = reinterpret_cast<T *>(p);
    p += seam_piece(sizeof(T) * count);
    return r;
  }
};
// A ladder: several systems of one model and precision stepped together,
// each MD kernel launched once for all of them (g6_ladder.hip).  The block
// maps list every block of every system: one after another (flat, the
// default), or interleaved over the eight XCDs so that every block of a
// system runs on one XCD (workgroup b runs on XCD b mod 8).
struct Ladder {
  std::vector<System *> sys;
  int model = 0, prec = 0;
  ModelTable mt{};
  PairK<float> pkf{};
  PairK<double> pkd{};
  MdRep *reps = nullptr;                         // [nsys], device
  int2 *map_pre = nullptr, *map_main = nullptr;  // [grid], device
  int grid_pre = 0, grid_main = 0, reg_bytes = 0;
  double *th_dev = nullptr;  // [nsys][G6_THERMO_LEN] gathered thermo records
  int *fl_dev = nullptr;     // [nsys] gathered error codes
  double *th_host = nullptr; // pinned copies of the two
  int *fl_host = nullptr;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0;
  int map_mode = G6_LADDER_MAP_FLAT;
};
void md_launch_pre_lad(Ladder &l, bool correct, bool predict, bool sites, int parity);
void md_launch_close_lad(Ladder &l, int mode, int parity, int src_parity);
void md_launch_main_lad(Ladder &l, int parity, int close_mode);
void md_launch_gather_lad(Ladder &l);  // thermo records and error codes
// Set a system's thermostat to temp after a replica exchange
// (Mixture.F90:1630-1653): velocities scaled by sqrt(temp / old temp),
// zeta0..zeta4 cleared.  Enqueued on the system's stream.
int md_set_temperature(System &s, double temp);

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

namespace g6 {
inline bool is_cubic(const System &s) { return s.Lb[0] == s.Lb[1] && s.Lb[0] == s.Lb[2]; }
// Box_Volume (BoxModule.F90:160-165): x y z
__device__ __host__ inline double box_volume(const double *L) { return (L[0] * L[1]) * L[2]; }
}  // namespace g6
