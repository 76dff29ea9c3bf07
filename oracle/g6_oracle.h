/*
 * CPU oracle for the rigid-water interaction step: a plain-C, FP64, serial
 * restatement of the Genesis6 reference algorithms, used ONLY as a checker
 * (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg).  It is
 * never linked into, loaded by, or called from the product library.
 *
 * Every function cites the reference file:line it restates.  Molecule and
 * pair indices are 0-based here (the reference is 1-based).  Arrays are flat
 * and caller-owned.
 */
#ifndef G6_ORACLE_H
#define G6_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { G6O_TIP4P = 0, G6O_TIP5P = 1, G6O_ST2 = 2, G6O_SPCE = 3, G6O_NVDE = 4, G6O_OPLSMEOH = 5 };
#define G6O_MAXSITE 7

/* Rigid-model tables: site charges/LJ, body-frame geometry, inertia, mass.
 * TIP4P: TIP4PModule.F90:43-97, TIP4PConst.F90:4-22
 * TIP5P: TIP5PModule.F90:39-83, RigidModule.F90:429-477
 * ST2:   ST2Module.F90:10-100
 * SPC/E: SPCEModule.F90:7-82 (StdForce, Interaction2.F90:67)
 * NvdE:  NvdEModule.F90:10-148 (StdForce)                                 */
typedef struct {
  int nsite;                                  /* including the CoM site   */
  double q[G6O_MAXSITE], eps[G6O_MAXSITE], sig[G6O_MAXSITE];
  double bx[G6O_MAXSITE], by[G6O_MAXSITE], bz[G6O_MAXSITE];
  double Ixx, Iyy, Izz, mass;
} g6o_model;
int g6o_model_get(int model, g6o_model *m);
/* Coulomb prefactors and LJ A,B in internal units, as the custom kernels
 * define them (TIP4PConst.F90:4-13, ST2Module.F90:121-125).               */
void g6o_constants(double *out16);

/* ---- grid and pair list (GridModule.F90, InteractionModule.F90) ------- */
/* ndiv = int(L/req) (GridModule.F90:487-501)                              */
int g6o_grid_ndiv(double L, double req);
/* half-shell stencil with mirror dedup, maxradius = 0 branch
 * (GridModule.F90:279-374); returns n, writes n0; offsets 0..ndiv-1.       */
int g6o_stencil_homo(const int ndiv[3], int *sx, int *sy, int *sz,
                     int *n0, int *sx0, int *sy0, int *sz0);
/* cell of each molecule (Address_Assign, GridModule.F90:503-552); 0-based
 * cell index mAddr-1.                                                       */
void g6o_address_assign(int n, const double *com, double L, const int ndiv[3],
                        int *cell);
/* Candidate list in reference order (Grid_NeighborList,
 * GridModule.F90:557-752) followed by the ordered FP64 compression
 * (Interaction_Compress2 with fbox, InteractionModule.F90:398-605).
 * Returns npair (or -1 if cap is exceeded); *npair0 = candidates.          */
long g6o_pairlist(int n, const double *com, double L, double rc, double req,
                  long cap, long *npair0, int *pi, int *pj,
                  double *ox, double *oy, double *oz);
/* all-pairs candidate list (interaction_alltoall, InteractionModule.F90:157-196)
 * followed by the same compression.                                        */
long g6o_pairlist_alltoall(int n, const double *com, double L, double rc,
                           long cap, int *pi, int *pj, double *ox, double *oy,
                           double *oz);
/* eratio/fratio per pair (CutOff_Dimmer2, CutoffModule.F90:220-278).      */
void g6o_cutoff_dimmer(long npair, const int *pi, const int *pj,
                       const double *com, const double *ox, const double *oy,
                       const double *oz, double rin, double rout,
                       double *eratio, double *fratio);
void g6o_cutoff_func(double rin, double rout, double dd, double *hh, double *dh);

/* ---- site kernels ------------------------------------------------------ */
/* Site arrays are SoA over n*nsite sites, molecule-major (site index
 * i*nsite + s), as sSite (SiteModule.F90:29-37).  Forces are ADDED.        */
/* interaction_force_tip4p (TIP4PModule.F90:187-785)                        */
void g6o_force_tip4p(long npair, const int *pi, const int *pj,
                     const double *ox, const double *oy, const double *oz,
                     const double *eratio, const double *fratio,
                     const double *x, const double *y, const double *z,
                     double *fx, double *fy, double *fz, double *ep,
                     double *vrmat);
/* StdForce (StdInteraction.F90:63-630), any model                          */
void g6o_force_std(int model, long npair, const int *pi, const int *pj,
                   const double *ox, const double *oy, const double *oz,
                   const double *eratio, const double *fratio,
                   const double *x, const double *y, const double *z,
                   double *fx, double *fy, double *fz, double *ep,
                   double *vrmat);
/* force_st2 with the Stillinger-Rahman switch (ST2Module.F90:112-897)      */
void g6o_force_st2(long npair, const int *pi, const int *pj,
                   const double *ox, const double *oy, const double *oz,
                   const double *eratio, const double *fratio,
                   const double *x, const double *y, const double *z,
                   double *fx, double *fy, double *fz, double *ep,
                   double *vrmat);
/* StdPairPotential energy only (StdInteraction.F90:633-762)               */
void g6o_energy_std(int model, long npair, const int *pi, const int *pj,
                    const double *ox, const double *oy, const double *oz,
                    const double *eratio, const double *x, const double *y,
                    const double *z, double *ep);

/* ---- rigid bodies (RigidModule.F90) ------------------------------------ */
/* state layout: com[3n] (x,y,z interleaved per molecule), quat[5][4][n],
 * dx[5][3][n] (Nordsieck, dx(1)=v*dt), w[5][3][n] (body omega, rad/ps)     */
void g6o_setsites(int model, int n, const double *com, const double *quat,
                  double *x, double *y, double *z, double *tmat, double *intra);
void g6o_collect(int model, int n, const double *fx, const double *fy,
                 const double *fz, const double *tmat, const double *intra,
                 double *force, double *torque);
void g6o_predict(int n, double *com, double *dx, double *w, double *quat);
void g6o_correct(int model, int n, double dt, double zeta0, const double *force,
                 const double *torque, double *com, double *dx, double *w,
                 double *quat);
void g6o_ek(int model, int n, double dt, const double *dx, const double *w,
            double *ekt, double *ekr);
/* Nose state: [temp, q(internal, x dt), dump, zeta0..zeta4, zetasum]      */
void g6o_correct_npt(int model, int n, double dt, double zeta0, double acoeff,
                     const double *force, const double *torque, double *com, double *dx,
                     double *w, double *quat);
void g6o_andersen_predict(double *a, double *L);
void g6o_andersen_correct(double *a, double *L, double pin);
void g6o_nose_predict(double *nose);
void g6o_nose_correct(double *nose, double temperature);

/* ---- MD driver: the Mixture.F90:1513-2286 step order ------------------ */
typedef struct {
  int model, n, fast, nose_active;
  double L, rc, taper, req, dt;
  double nose[9];
  double *com, *dx, *w, *quat;      /* state, as above                    */
  double ep, ekt, ekr, temperature, vir[9];
  long npair;
  /* Andersen barostat, orthorhombic (cube): anders = pex (atm), mass
   * (internal: @PCN4 mass * dof / dt^2, Mixture.F90:1207-1231), v0..v5 */
  int andersen;
  double anders[8];
  double pressure, volume, andersen_ep, andersen_ek; /* atm, A^3, J/mol */
} g6o_md;
/* one step; returns npair or -1 */
long g6o_md_step(g6o_md *s);
/* Conserved quantity of the @LOGD line (Mixture.F90:2431-2468), J/mol.    */
double g6o_md_hamiltonian(const g6o_md *s);

/* ---- MT19937 and the reference's velocity initialisation ------------- */
void g6o_mt_seed(uint32_t seed);
uint32_t g6o_mt_next(void);
/* Random_GetNext: 1 - (mt>>1)/2^31 (RandomModule.F90:123-130, mtrnd.f)    */
double g6o_random_getnext(void);
/* Rigid_VReset2 (RigidModule.F90:1016-1053): writes dx(1), w(1).          */
void g6o_vreset(int model, int n, double dt, double temp, uint32_t seed,
                double *dx, double *w);

/* Monte Carlo (mc2.F90:303-340, 540-862; PairInfo.F90:12-46).
 * com[n][3], quat[4][n] (order-1 quaternions), ep[n*n] the PairInfo matrix. */
double g6o_mc_pairinfo(int model, int n, double L, double rc, double taper,
                       const double *com, const double *quat, double *ep);
long g6o_mc_trials(int model, int n, double L, double rc, double taper, double temp,
                   double *com, double *quat, double *delta, double *rtr, double *ep,
                   double *epsum, const double *u, long *pos, int ntrial);

/* The trial loop with umbrella potentials coeffi (op - center)^2
 * (Umbrella.F90:11-16) added to every trial's energy change (mc2.F90:684-706):
 * type 1 TRIANGLE, 4 LOGTRIAN (sTriangle, Triangle.F90:53-280, on refhist[21]),
 * 3 EPOP____ (op = the energy).  One call is one mc2 loop: the order
 * parameters are prepared on the loop's pair list first (mc2.F90:486-521).
 * op[numb] returns each order parameter after the last trial.  Returns the
 * accepted count, or -1 (an unknown type, or a neighbour array overflow). */
#define G6O_MAXUMB 10
long g6o_mc_trials_umbrella(int model, int n, double L, double rc, double taper, double temp,
                            double *com, double *quat, double *delta, double *rtr, double *ep,
                            double *epsum, const double *u, long *pos, int ntrial, int numb,
                            const int *type, const double *center, const double *coeffi,
                            const double *refhist, double *op);

/* NPT Monte Carlo with collective moves (no reference counterpart; the
 * device path's definition, restated serially) */
void g6o_philox10(uint32_t *c, uint32_t k0, uint32_t k1);
void g6o_npt_run(int model, int n, double rc, double taper, double voxa, double temp,
                 double press_bar, uint64_t seed, double disp_max, double rot_max_deg,
                 double dlnv_max, double p_volume, double *L, double *com, double *quat,
                 double *U, long long *index, int nmoves, long long *counts, int *accepted);

/* Neighbour arrays (Neighbor.F90).  A table of n arrays: nnear[n] and
 * near/dx/dy/dz/dd[n][G6O_NEAR] (G6O_NEAR = MAXNEAR+1 = 151 entries per
 * array), 0-based labels.  radius <= 0 means the argument is absent.
 * Return 0, or 1 when an array would exceed G6O_NEAR entries (the
 * reference then writes past its arrays). */
#define G6O_NEAR 151
typedef struct {
  int *nnear, *near;
  double *dx, *dy, *dz, *dd;
} g6o_nbtab;
/* neighbors_all2 (Neighbor.F90:91-159) over a pair list */
int g6o_neighbors_all2(int nsite, long npair, const int *pi, const int *pj, const double *ox,
                       const double *oy, const double *oz, const double *coord, int maxnei,
                       int order, double radius, g6o_nbtab out);
/* neighbor1_compress2 (Neighbor.F90:195-215) applied to arrays 0..nrow-1 */
int g6o_neighbors_compress(int nrow, g6o_nbtab in, int maxnei, int order, double radius,
                           g6o_nbtab out);
/* neighbors_displacement (Neighbor.F90:401-457): *nmod arrays and their
 * molecule indices idx[] */
int g6o_neighbors_displacement(g6o_nbtab in, int order, int target, double ddx, double ddy,
                               double ddz, int *nmod, int *idx, g6o_nbtab out);

/* Umbrella order parameters of mc2 on the neighbour arrays: Q6 local bond
 * order (Q6Local.F90:101-231, q6.F90) and the triangle histogram
 * (Triangle.F90:53-116, DistribModule.F90:149-158, 304-326) */
void g6o_q6_qlm(int m, double x, double y, double z, double *re, double *im);
int g6o_q6local(int nsite, long npair, const int *pi, const int *pj, const double *ox,
                const double *oy, const double *oz, const double *coord, double *q6,
                int *nbond, double *sumq6, double *op);
int g6o_triangle(int nsite, long npair, const int *pi, const int *pj, const double *ox,
                 const double *oy, const double *oz, const double *coord, const double *refhist,
                 int logarithmic, double *hist, double *op);

#ifdef __cplusplus
}
#endif

#endif
