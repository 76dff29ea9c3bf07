// Replica ladders and replica exchange.
//
// The reference runs replica-exchange MD as one fakempi process per replica:
// every node steps its own system @INNR steps, node 0 gathers (epsum, beta),
// tries @NRXG swaps over the exchange graph (replica_exchange_trial,
// Replica.F90:7-73) and scatters the betas, and each node whose beta changed
// rescales its velocities and resets its thermostat (Mixture.F90:1570-1653).
//
// Here the replicas are systems of one process and one device.  A ladder
// steps all of them with one launch per MD kernel (k_md_*_lad, g6_md.hip):
// the block maps below list every block of every system (one after another,
// or interleaved over the eight XCDs so that each system stays on one XCD;
// the autotuner times both), and the launches of a run of steps are
// captured into one hipGraph.  An exchange group holds the
// node-0 state (MT19937 stream, node_in_charge) and applies the swap rule on
// the host: it reads 64 numbers per exchange, so it does not belong on the
// device.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "g6_internal.h"

#pragma clang fp contract(off)

namespace g6 {

constexpr double kI2J_rx = 10.0, kJ2K_rx = 0.120273;  // PhysConst.F90:48, 56

// thermo record and error code of every system into two dense arrays
__global__ __launch_bounds__(64) void k_lad_gather(const MdRep *__restrict__ reps, double *th,
                                                   int *fl) {
  const MdRep &r = reps[blockIdx.x];
  if (threadIdx.x < G6_THERMO_LEN)
    th[(size_t)blockIdx.x * G6_THERMO_LEN + threadIdx.x] = r.b.thermo[threadIdx.x];
  if (threadIdx.x == 0) fl[blockIdx.x] = r.b.flags[0];
}

// Rigid_ScaleVelocity (RigidModule.F90:621-633): dx(1) and w(1) times ratio;
// then the thermostat retarget of Mixture.F90:1644-1649
__global__ __launch_bounds__(256) void k_set_temperature(double *dx, double *w, double *nose,
                                                         double *thermo, int n, double ratio,
                                                         double temp) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < 3 * n) {
    dx[k] = dx[k] * ratio;
    w[k] = w[k] * ratio;
  }
  if (k == 0) {
    nose[0] = temp;
    for (int z = 3; z < 8; z++) nose[z] = 0.0;
    thermo[29] = temp;
  }
}

void md_launch_gather_lad(Ladder &l) {
  k_lad_gather<<<(int)l.sys.size(), 64, 0, l.stream>>>(l.reps, l.th_dev, l.fl_dev);
}

static int set_temperature_from(System &s, double temp, double old_temp) {
  if (!s.nose_active || !(temp > 0) || !(old_temp > 0)) return G6_EARG;
  const double ratio = sqrt(temp / old_temp);
  k_set_temperature<<<(3 * s.n + 255) / 256, 256, 0, s.stream>>>(s.md.dx, s.md.w, s.nose,
                                                                 s.thermo, s.n, ratio, temp);
  G6_CHECK(hipGetLastError());
  return 0;
}

// k steps of every system, the launch sequence of g6_step's launch_steps
static void launch_steps_lad(Ladder &l, int k) {
  for (int q = 0; q < k; q++) {
    const int p = q & 1;
    md_launch_pre_lad(l, q > 0, true, true, p);
    if (q == 0) md_launch_close_lad(l, 2 | 4, p, 1 - p);
    md_launch_main_lad(l, p, q > 0 ? (1 | 2 | 4) : 0);
  }
  md_launch_pre_lad(l, true, false, false, 0);
  md_launch_close_lad(l, 1 | 8, 0, (k - 1) & 1);
}

// Block maps: systems go to the XCD lane with the fewest molecules so far
// (largest first); lane x's blocks take workgroup slots x, x + 8, x + 16, ...
// and short lanes are padded with idle blocks.
// first: the block index each system's list starts from (-1 adds the close
// block of k_md_main_lad)
static void build_map(const std::vector<int> &lane_of, const std::vector<int> &nblk, int first,
                      std::vector<int2> &map) {
  std::vector<std::vector<int2>> lanes(8);
  for (size_t r = 0; r < nblk.size(); r++)
    for (int b = first; b < nblk[r]; b++) lanes[lane_of[r]].push_back(make_int2((int)r, b));
  size_t len = 0;
  for (auto &v : lanes) len = std::max(len, v.size());
  map.assign(8 * len, make_int2(-1, 0));
  for (int x = 0; x < 8; x++)
    for (size_t k = 0; k < lanes[x].size(); k++) map[8 * k + x] = lanes[x][k];
}

static int ladder_steps_setting() {
  const char *e = getenv("G6_GRAPH_STEPS");
  return e && atoi(e) <= 0 ? 0 : 1000;  // 0 disables graphs, as for g6_step
}

struct Rx {
  int nrep = 0;
  std::vector<System *> sys;
  std::vector<int> a, b;  // bonds, 1-based nodes
  int nexchange = 0;
  McHost rand;
  std::vector<int> nic;  // node_in_charge, 1-based
};

static std::mutex g_lmu;
static std::vector<Ladder *> g_lad;
static std::vector<Rx *> g_rx;

static Ladder *lget(g6_ladder h) {
  std::lock_guard<std::mutex> lk(g_lmu);
  if (h <= 0 || h > (g6_ladder)g_lad.size()) return nullptr;
  return g_lad[h - 1];
}
static Rx *xget(g6_rx h) {
  std::lock_guard<std::mutex> lk(g_lmu);
  if (h <= 0 || h > (g6_rx)g_rx.size()) return nullptr;
  return g_rx[h - 1];
}

static void ladder_release(Ladder &l) {
  if (l.graph) (void)hipGraphExecDestroy(l.graph);
  if (l.reps) (void)hipFree(l.reps);
  if (l.map_pre) (void)hipFree(l.map_pre);
  if (l.map_main) (void)hipFree(l.map_main);
  if (l.th_dev) (void)hipFree(l.th_dev);
  if (l.fl_dev) (void)hipFree(l.fl_dev);
  if (l.th_host) (void)hipHostFree(l.th_host);
  if (l.fl_host) (void)hipHostFree(l.fl_host);
  if (l.stream) (void)hipStreamDestroy(l.stream);
}

static std::vector<MdRep> descriptors(const Ladder &l) {
  std::vector<MdRep> d(l.sys.size());
  for (size_t r = 0; r < l.sys.size(); r++) {
    const System &s = *l.sys[r];
    d[r].g = s.geo_md;
    d[r].b = s.md;
    d[r].dt = s.dt;
    d[r].nose_active = s.nose_active ? 1 : 0;
    d[r].nblk_main = s.nblk_md;
  }
  return d;
}

// contiguous map: the systems' blocks one after another (a system spread
// over every XCD)
static void build_map_flat(const std::vector<int> &nblk, int first, std::vector<int2> &map) {
  map.clear();
  for (size_t r = 0; r < nblk.size(); r++)
    for (int b = first; b < nblk[r]; b++) map.push_back(make_int2((int)r, b));
}

static int ladder_maps(Ladder &l) {
  const size_t ns = l.sys.size();
  std::vector<int> order(ns), lane_of(ns), np(ns), nm(ns);
  for (size_t r = 0; r < ns; r++) {
    order[r] = (int)r;
    np[r] = pre_waves(l.sys[r]->n) / 4;
    nm[r] = l.sys[r]->nblk_md;
  }
  std::vector<int2> mp, mm;
  if (l.map_mode == G6_LADDER_MAP_XCD) {
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return l.sys[x]->n > l.sys[y]->n; });
    long long load[8] = {0};
    for (int r : order) {
      int best = 0;
      for (int x = 1; x < 8; x++)
        if (load[x] < load[best]) best = x;
      lane_of[r] = best;
      load[best] += l.sys[r]->n;
    }
    build_map(lane_of, np, 0, mp);
    build_map(lane_of, nm, -1, mm);
  } else {
    build_map_flat(np, 0, mp);
    build_map_flat(nm, -1, mm);
  }
  if (l.map_pre) (void)hipFree(l.map_pre);
  if (l.map_main) (void)hipFree(l.map_main);
  l.map_pre = l.map_main = nullptr;
  l.grid_pre = (int)mp.size();
  l.grid_main = (int)mm.size();
  G6_CHECK(hipMalloc(&l.map_pre, sizeof(int2) * mp.size()));
  G6_CHECK(hipMalloc(&l.map_main, sizeof(int2) * mm.size()));
  G6_CHECK(hipMemcpy(l.map_pre, mp.data(), sizeof(int2) * mp.size(), hipMemcpyHostToDevice));
  G6_CHECK(hipMemcpy(l.map_main, mm.data(), sizeof(int2) * mm.size(), hipMemcpyHostToDevice));
  if (l.graph) (void)hipGraphExecDestroy(l.graph);  // captured with the old maps
  l.graph = nullptr;
  return 0;
}

static int ladder_setup(Ladder &l) {
  const size_t ns = l.sys.size();
  const int rm = ladder_maps(l);
  if (rm) return rm;
  int rb = 0;
  for (System *s : l.sys)
    rb = std::max(rb, main_region_bytes(s->md.maxnb, l.prec == G6_FP32 ? 4 : 8));
  l.reg_bytes = rb;
  const std::vector<MdRep> d = descriptors(l);
  G6_CHECK(hipMalloc(&l.reps, sizeof(MdRep) * ns));
  G6_CHECK(hipMalloc(&l.th_dev, sizeof(double) * G6_THERMO_LEN * ns));
  G6_CHECK(hipMalloc(&l.fl_dev, sizeof(int) * ns));
  G6_CHECK(hipHostMalloc(&l.th_host, sizeof(double) * G6_THERMO_LEN * ns, hipHostMallocDefault));
  G6_CHECK(hipHostMalloc(&l.fl_host, sizeof(int) * ns, hipHostMallocDefault));
  memset(l.th_host, 0, sizeof(double) * G6_THERMO_LEN * ns);
  G6_CHECK(hipMemcpy(l.reps, d.data(), sizeof(MdRep) * ns, hipMemcpyHostToDevice));
  G6_CHECK(hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking));
  return 0;
}

static double rx_umbrella(double coeffi, double center, double op) {
  return coeffi * ((op - center) * (op - center));  // Umbrella.F90:11-16
}

// replica_exchange_trial (Replica.F90:7-73) inside the exchange loop of
// Mixture.F90:1605-1622 (numb = 0) or of mc2.F90:973-996: with numb
// umbrellas per node, each state's energy adds its umbrellas at its order
// parameters (op, center, coeffi indexed (node-1)*numb + k), the swapped
// state the other node's umbrellas, and an accepted swap exchanges the
// nodes' umbrellas with their betas
static int rx_trials(Rx &x, const double *energy, double *beta, int numb = 0,
                     const double *op = nullptr, double *center = nullptr,
                     double *coeffi = nullptr) {
  int acc = 0;
  const int nbond = (int)x.a.size();
  for (int t = 0; t < x.nexchange; t++) {
    // k = random_getnext * nbond + 1, truncated; Random_GetNext can return
    // exactly 1, which would index one past the last bond
    int k = (int)(x.rand.getnext() * nbond + 1.0);
    if (k > nbond) k = nbond;
    const int j = x.a[k - 1], jj = x.b[k - 1];
    const int n0 = x.nic[j - 1], n1 = x.nic[jj - 1];
    const int n00 = (n0 - 1) * numb, n11 = (n1 - 1) * numb;
    double a0 = energy[n0 - 1], a1 = energy[n1 - 1];
    for (int u = 0; u < numb; u++) {
      a0 = a0 + rx_umbrella(coeffi[n00 + u], center[n00 + u], op[n00 + u]);
      a1 = a1 + rx_umbrella(coeffi[n11 + u], center[n11 + u], op[n11 + u]);
    }
    const double acc0 = -(a0 * beta[n0 - 1] + a1 * beta[n1 - 1]) * kI2J_rx * kJ2K_rx;
    a0 = energy[n0 - 1];
    a1 = energy[n1 - 1];
    for (int u = 0; u < numb; u++) {
      a0 = a0 + rx_umbrella(coeffi[n11 + u], center[n11 + u], op[n00 + u]);
      a1 = a1 + rx_umbrella(coeffi[n00 + u], center[n00 + u], op[n11 + u]);
    }
    const double acc1 = -(a0 * beta[n1 - 1] + a1 * beta[n0 - 1]) * kI2J_rx * kJ2K_rx;
    bool accept = false;
    if (acc0 < acc1) {
      accept = true;
    } else {
      const double ratio = exp(acc1 - acc0);
      if (x.rand.getnext() < ratio) accept = true;
    }
    if (accept) {
      for (int u = 0; u < numb; u++) {
        std::swap(center[n00 + u], center[n11 + u]);
        std::swap(coeffi[n00 + u], coeffi[n11 + u]);
      }
      const double d = beta[n0 - 1];
      beta[n0 - 1] = beta[n1 - 1];
      beta[n1 - 1] = d;
      x.nic[j - 1] = n1;
      x.nic[jj - 1] = n0;
      acc++;
    }
  }
  return acc;
}

}  // namespace g6

using namespace g6;

extern "C" {

int g6_ladder_create(g6_ladder *hl, int nsys, const g6_handle *hs) {
  if (!hl || nsys < 1 || !hs) return G6_EARG;
  Ladder *l = new Ladder();
  for (int r = 0; r < nsys; r++) {
    System *s = system_of(hs[r]);
    bool dup = false;
    for (System *t : l->sys) dup = dup || t == s;
    if (!s || dup || s->andersen || !is_cubic(*s)) {  // ladders step fixed cubic boxes
      delete l;
      return G6_EARG;
    }
    if (r == 0) {
      l->model = s->kmodel;  // the pair-kernel layout
      l->prec = s->prec;
      l->mt = s->mt;
      l->pkf = s->pkf;
      l->pkd = s->pkd;
    } else if (s->kmodel != l->model || s->prec != l->prec ||
               memcmp(&s->pkf, &l->pkf, sizeof(l->pkf)) != 0 ||
               memcmp(&s->pkd, &l->pkd, sizeof(l->pkd)) != 0) {
      delete l;  // one kernel instance serves every system
      return G6_EARG;
    }
    l->sys.push_back(s);
  }
  // the ladder kernels run the rows of both directions
  for (System *s : l->sys) {
    const int rg = use_row_grid(*s);
    if (rg != 0) {
      delete l;
      return rg;
    }
    s->tile_ok = false;  // a ladder's systems keep the rows
  }
  const int rc = ladder_setup(*l);
  if (rc != 0) {
    ladder_release(*l);
    delete l;
    return rc;
  }
  std::lock_guard<std::mutex> lk(g_lmu);
  g_lad.push_back(l);
  *hl = (g6_ladder)g_lad.size();
  return 0;
}

int g6_ladder_free(g6_ladder hl) {
  Ladder *l = lget(hl);
  if (!l) return G6_EARG;
  (void)hipStreamSynchronize(l->stream);
  ladder_release(*l);
  delete l;
  std::lock_guard<std::mutex> lk(g_lmu);
  g_lad[hl - 1] = nullptr;
  return 0;
}

int g6_ladder_step(g6_ladder hl, int nsteps) {
  Ladder *l = lget(hl);
  if (!l || nsteps < 0) return G6_EARG;
  for (System *s : l->sys) G6_CHECK(hipStreamSynchronize(s->stream));
  // the descriptors carry each system's thermostat switch
  const std::vector<MdRep> d = descriptors(*l);
  G6_CHECK(hipMemcpy(l->reps, d.data(), sizeof(MdRep) * d.size(), hipMemcpyHostToDevice));
  if (nsteps > 0) {
    const int G = ladder_steps_setting() ? std::min(nsteps, 1000) : 0;
    if (G > 0) {
      if (!l->graph || l->graph_steps != G) {
        if (l->graph) {
          (void)hipGraphExecDestroy(l->graph);
          l->graph = nullptr;
        }
        hipGraph_t gr;
        G6_CHECK(hipStreamBeginCapture(l->stream, hipStreamCaptureModeThreadLocal));
        launch_steps_lad(*l, G);
        G6_CHECK(hipStreamEndCapture(l->stream, &gr));
        G6_CHECK(hipGraphInstantiate(&l->graph, gr, nullptr, nullptr, 0));
        G6_CHECK(hipGraphDestroy(gr));
        l->graph_steps = G;
      }
      for (int k = 0; k + G <= nsteps; k += G) G6_CHECK(hipGraphLaunch(l->graph, l->stream));
    }
    const int rest = G > 0 ? nsteps % G : nsteps;
    if (rest > 0) launch_steps_lad(*l, rest);
  }
  md_launch_gather_lad(*l);
  G6_CHECK(hipGetLastError());
  const size_t ns = l->sys.size();
  G6_CHECK(hipMemcpyAsync(l->th_host, l->th_dev, sizeof(double) * G6_THERMO_LEN * ns,
                          hipMemcpyDeviceToHost, l->stream));
  G6_CHECK(hipMemcpyAsync(l->fl_host, l->fl_dev, sizeof(int) * ns, hipMemcpyDeviceToHost,
                          l->stream));
  G6_CHECK(hipStreamSynchronize(l->stream));
  int rc = 0;
  for (size_t r = 0; r < ns; r++) {
    l->sys[r]->steps_done += nsteps;
    const int f = l->fl_host[r];
    if (f == 1 || f == 3 || f == 4) rc = G6_ECAPACITY;
    else if (f == 2 && rc == 0) rc = G6_EARG;
  }
  return rc;
}

int g6_ladder_time_force(g6_ladder hl, int nlaunch, double *ms) {
  Ladder *l = lget(hl);
  if (!l || nlaunch < 1 || nlaunch > 10000 || !ms) return G6_EARG;
  // the ladder's force kernel alone (every system's blocks in one launch),
  // nlaunch times back to back on the state a g6_ladder_step call leaves;
  // one graph, or eager launches when G6_GRAPH_STEPS=0
  for (System *s : l->sys) G6_CHECK(hipStreamSynchronize(s->stream));
  G6_CHECK(hipStreamSynchronize(l->stream));
  hipEvent_t e0, e1;
  G6_CHECK(hipEventCreate(&e0));
  const hipError_t ec = hipEventCreate(&e1);
  if (ec != hipSuccess) {
    (void)hipEventDestroy(e0);
    return G6_EHIP;
  }
  hipGraphExec_t ge = nullptr;
  hipError_t err = hipSuccess;
  const bool graph = ladder_steps_setting() > 0;
  if (graph) {
    hipGraph_t gr;
    err = hipStreamBeginCapture(l->stream, hipStreamCaptureModeThreadLocal);
    if (err == hipSuccess) {
      for (int k = 0; k < nlaunch; k++) md_launch_main_lad(*l, 0, 0);
      err = hipStreamEndCapture(l->stream, &gr);
      if (err == hipSuccess) {
        err = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
      }
    }
    if (err == hipSuccess) err = hipGraphLaunch(ge, l->stream);  // warm
  } else {
    md_launch_main_lad(*l, 0, 0);
  }
  if (err == hipSuccess) err = hipEventRecord(e0, l->stream);
  if (err == hipSuccess) {
    if (graph) {
      err = hipGraphLaunch(ge, l->stream);
    } else {
      for (int k = 0; k < nlaunch; k++) md_launch_main_lad(*l, 0, 0);
      err = hipGetLastError();
    }
  }
  if (err == hipSuccess) err = hipEventRecord(e1, l->stream);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  float e = 0;
  if (err == hipSuccess) err = hipEventElapsedTime(&e, e0, e1);
  if (ge) (void)hipGraphExecDestroy(ge);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (err != hipSuccess) return G6_EHIP;
  *ms = e / nlaunch;
  return 0;
}

int g6_ladder_set_map(g6_ladder hl, int mode) {
  Ladder *l = lget(hl);
  if (!l || (mode != G6_LADDER_MAP_XCD && mode != G6_LADDER_MAP_FLAT)) return G6_EARG;
  G6_CHECK(hipStreamSynchronize(l->stream));
  l->map_mode = mode;
  return ladder_maps(*l);
}

int g6_ladder_thermo(g6_ladder hl, double *thermo) {
  Ladder *l = lget(hl);
  if (!l || !thermo) return G6_EARG;
  memcpy(thermo, l->th_host, sizeof(double) * G6_THERMO_LEN * l->sys.size());
  return 0;
}

int g6_set_temperature(g6_handle h, double temp) {
  System *s = system_of(h);
  if (!s) return G6_EARG;
  double z[9];
  G6_CHECK(hipStreamSynchronize(s->stream));
  G6_CHECK(hipMemcpy(z, s->nose, sizeof(z), hipMemcpyDeviceToHost));
  const int rc = set_temperature_from(*s, temp, z[0]);
  if (rc != 0) return rc;
  G6_CHECK(hipStreamSynchronize(s->stream));
  return 0;
}

int g6_rx_create(g6_rx *hx, int nrep, const g6_handle *hs, int nbond, const int32_t *a,
                 const int32_t *b, int nexchange, uint32_t seed) {
  if (!hx || nrep < 1 || nbond < 1 || !a || !b || nexchange < 0) return G6_EARG;
  Rx *x = new Rx();
  x->nrep = nrep;
  for (int k = 0; k < nbond; k++) {
    if (a[k] < 1 || a[k] > nrep || b[k] < 1 || b[k] > nrep || a[k] == b[k]) {
      delete x;
      return G6_EARG;
    }
    x->a.push_back(a[k]);
    x->b.push_back(b[k]);
  }
  if (hs) {
    for (int r = 0; r < nrep; r++) {
      System *s = system_of(hs[r]);
      if (!s) {
        delete x;
        return G6_EARG;
      }
      x->sys.push_back(s);
    }
  }
  x->nexchange = nexchange;
  x->rand.seed(seed);
  for (int r = 1; r <= nrep; r++) x->nic.push_back(r);
  std::lock_guard<std::mutex> lk(g_lmu);
  g_rx.push_back(x);
  *hx = (g6_rx)g_rx.size();
  return 0;
}

int g6_rx_free(g6_rx hx) {
  Rx *x = xget(hx);
  if (!x) return G6_EARG;
  delete x;
  std::lock_guard<std::mutex> lk(g_lmu);
  g_rx[hx - 1] = nullptr;
  return 0;
}

int g6_rx_trials(g6_rx hx, const double *energy, double *beta, int32_t *naccept) {
  Rx *x = xget(hx);
  if (!x || !energy || !beta) return G6_EARG;
  const int acc = rx_trials(*x, energy, beta);
  if (naccept) *naccept = acc;
  return 0;
}

int g6_rx_trials_umbrella(g6_rx hx, const double *energy, double *beta, int numbrella,
                          const double *op, double *center, double *coeffi,
                          int32_t *naccept) {
  Rx *x = xget(hx);
  if (!x || !energy || !beta || numbrella < 0) return G6_EARG;
  if (numbrella > 0 && (!op || !center || !coeffi)) return G6_EARG;
  const int acc = rx_trials(*x, energy, beta, numbrella, op, center, coeffi);
  if (naccept) *naccept = acc;
  return 0;
}

int g6_rx_nodes(g6_rx hx, int32_t *node_in_charge) {
  Rx *x = xget(hx);
  if (!x || !node_in_charge) return G6_EARG;
  for (int r = 0; r < x->nrep; r++) node_in_charge[r] = x->nic[r];
  return 0;
}

int g6_rx_exchange(g6_rx hx, const double *thermo, int32_t *naccept) {
  Rx *x = xget(hx);
  if (!x || (int)x->sys.size() != x->nrep) return G6_EARG;
  const int nr = x->nrep;
  std::vector<double> th((size_t)nr * G6_THERMO_LEN);
  if (thermo) {
    memcpy(th.data(), thermo, sizeof(double) * th.size());
  } else {
    for (int r = 0; r < nr; r++) {
      System &s = *x->sys[r];
      G6_CHECK(hipStreamSynchronize(s.stream));
      G6_CHECK(hipMemcpy(&th[(size_t)r * G6_THERMO_LEN], s.thermo,
                         sizeof(double) * G6_THERMO_LEN, hipMemcpyDeviceToHost));
    }
  }
  std::vector<double> energy(nr), beta(nr), beta0(nr), told(nr);
  for (int r = 0; r < nr; r++) {
    if (!x->sys[r]->nose_active) return G6_EARG;  // no thermostat to retarget
    told[r] = th[(size_t)r * G6_THERMO_LEN + 29];  // mdreplica%temp = nose%temp
    if (!(told[r] > 0)) return G6_EARG;
    energy[r] = th[(size_t)r * G6_THERMO_LEN + 0];  // epsum
    beta[r] = 1.0 / told[r];
    beta0[r] = beta[r];
  }
  const int acc = rx_trials(*x, energy.data(), beta.data());
  for (int r = 0; r < nr; r++) {
    if (beta[r] != beta0[r]) {
      const int rc = set_temperature_from(*x->sys[r], 1.0 / beta[r], told[r]);
      if (rc != 0) return rc;
    }
  }
  for (int r = 0; r < nr; r++) G6_CHECK(hipStreamSynchronize(x->sys[r]->stream));
  if (naccept) *naccept = acc;
  return 0;
}

}  // extern "C"
