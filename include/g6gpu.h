/*
 * g6gpu: C ABI of the MI355X rigid-water interaction step.
 *
 * The drop-in seam is the reference's Fortran module interface for the MD
 * step (Mixture.F90:1789-2286) and for the MC trial energies (mc2.F90:544-862).
 * Each entry point below names the reference routine it replaces.  Calling
 * conventions (so that ISO_C_BINDING interfaces bind them directly):
 *   - every function returns int32: 0 = error_none; > 0 = the reference's
 *     error codes (Error.F90:2-7); < 0 = GPU/HIP or capacity errors (G6_E*);
 *   - scalars by value, arrays as plain host pointers to contiguous memory;
 *     no struct, no Fortran derived type crosses the boundary;
 *   - molecule/pair indices are 0-based; Fortran callers add 1;
 *   - per-molecule arrays are component-major ("SoA"): com[c*n + i] is
 *     component c of molecule i, i.e. Fortran com(n,3); Nordsieck arrays are
 *     [order][component][molecule], i.e. Fortran dx(n,3,5), quat(n,4,5);
 *   - all calls are synchronous: results are in the caller's arrays on return.
 * Units are the reference's internal units (A, ps, amu; energy 10 J/mol).
 */
#ifndef G6GPU_H
#define G6GPU_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t g6_handle;

/* models: the reference ids "TIP4P   ", "TIP5P   ", "ST2_____", "SPC_E   ",
 * "NVDE____" (RigidSetup.F90:33-62) */
enum { G6_TIP4P = 0, G6_TIP5P = 1, G6_ST2 = 2, G6_SPCE = 3, G6_NVDE = 4, G6_OPLSMEOH = 5 };
/* pair arithmetic precision; accumulation of forces, torques, energy and
 * virial is FP64 in both */
enum { G6_FP64 = 0, G6_FP32 = 1 };
/* error codes (< 0) */
enum {
  G6_EHIP = -1,       /* a HIP runtime call failed                        */
  G6_EARG = -2,       /* invalid argument (size, model, handle)           */
  G6_ECAPACITY = -3,  /* pair capacity exceeded: nothing truncated        */
  G6_ENODEVICE = -4   /* no HIP device                                    */
};
/* thermo record of the last step, written by g6_get_thermo (internal
 * energy units unless stated):
 * [0] ep (sum of pair energies), [1] ekt, [2] ekr, [3] temperature (K),
 * [4..12] pair virial vrmat(3,3) row-major, [13] conserved H (J/mol,
 * Mixture.F90:2431-2468), [14] interacting pairs (dd < rc^2),
 * [15] pairs in the current list (differs from [14] only when skin > 0),
 * [16..24] kinetic tensor of Rigid_PressureTensor (RigidModule.F90:600-619),
 * [25] pressure (atm, Mixture.F90:2243-2255), [26] volume (A^3),
 * [27] Nose potential term nose_ep, [28] Nose kinetic term nose_ek (J/mol,
 * Mixture.F90:2431-2434), [29] Nose target temperature nose%temp (K; 0
 * without a thermostat), [30] Andersen andersen_ep and [31] andersen_ek
 * (J/mol, Mixture.F90:2436-2439; 0 without a barostat) */
#define G6_THERMO_LEN 32

/* ---- system lifecycle (Mixture.F90:1141-1347 setup) -------------------- */
/* model: G6_*; n molecules in a cubic box of edge box_l (@BXLA); cutoff rc
 * with switch width taper (@RCOA rc taper); voxa = requested cell edge
 * (@VOXA); dt in ps (@DTPS); prec: G6_FP64/G6_FP32.  skin >= 0: pair-list
 * margin in A; 0 rebuilds the reference list every step (the reference's
 * @BOOK interval 1); > 0 rebuilds it when any molecule moved more than
 * skin/2 and filters dd < rc^2 in FP64 every step (same interacting set). */
int g6_init(g6_handle *h, int model, int n, double box_l, double rc,
            double taper, double voxa, double dt, int prec, double skin);
/* g6_init with an orthorhombic box: edges box3[3] (@BOX3, BoxModule.F90:
 * 109-116) and a requested cell edge per axis voxa3[3] (@VOXA x y z; the
 * reference grid divides each axis on its own, GridModule.F90:494-499).  The
 * MD step, the seam entry points and both barostat modes take such boxes;
 * the tile kernel, ladders, Monte Carlo and NPT Monte Carlo take cubic boxes
 * (G6_EARG otherwise; g6_step runs the rows of both directions). */
int g6_init_box3(g6_handle *h, int model, int n, const double *box3, double rc,
                 double taper, const double *voxa3, double dt, int prec, double skin);
int g6_finalize(g6_handle h);

/* Rigid state (RigidModule.F90:14-33 sRigid1, internal units).  com[3n];
 * quat[20n] = quat(n,4,5); dx[15n] = (dx,dy,dz)(n,3,5) Nordsieck, dx(1)=v*dt;
 * w[15n] = (wx,wy,wz)(n,3,5), body-frame omega in rad/ps. */
int g6_upload_state(g6_handle h, const double *com, const double *quat,
                    const double *dx, const double *w);
int g6_download_state(g6_handle h, double *com, double *quat, double *dx,
                      double *w);

/* The reference driver's @FAST switch (Mixture.F90:732-736, the custom
 * argument of force, Interaction2.F90:19-67).  fast = 0: every pair through
 * StdForce's coefficients (the reference's default), which for ST2 means
 * no Stillinger-Rahman switch; fast != 0: the custom TIP4P and ST2 kernels
 * (interaction_force_tip4p, force_st2).  Other models have one path.  The
 * library starts with fast = 1; the Monte Carlo paths always use
 * StdPairPotential, as mc2 does (mc2.F90:338). */
int g6_set_fast(g6_handle h, int fast);

/* Nose thermostat (NoseModule.F90:14-21).  nose[9] = temp (K), q (internal,
 * i.e. the @NOS2 value times dt, Mixture.F90:1206), dump, zeta0..zeta4,
 * zetasum.  active = 0 switches it off (NVE). */
int g6_set_nose(g6_handle h, int active, const double *nose);
int g6_get_nose(g6_handle h, double *nose);

/* Maxwell-Boltzmann velocities, Rigid_VReset2 (RigidModule.F90:1016-1053)
 * with the reference's MT19937 stream (@VSET temp seed). */
int g6_vreset(g6_handle h, double temp, uint32_t seed);

/* nsteps MD steps in the reference order (Mixture.F90:1789-2286): predict,
 * Nose_Predict, site placement, grid + pair list, cutoff ratios, forces,
 * collect, correct, Ek, Nose_Correct.  Replaces that loop body. */
int g6_step(g6_handle h, int nsteps);
int g6_get_thermo(g6_handle h, double *thermo);
/* Forces of the current state through the MD kernels, without moving it:
 * site placement, pair list, forces and rigid_collectforce
 * (RigidModule.F90:275-317).  force[3n] lab frame, torque[3n] body frame;
 * thermo as g6_get_thermo with ep, virial and pair counts of this state. */
int g6_eval_forces(g6_handle h, double *force, double *torque, double *thermo);
/* g6_eval_forces for callers that cannot allocate (no counterpart in the
 * reference): the same kernels and results, with the device staging buffer
 * of force and torque taken from a fixed device array of 16 MiB that each
 * GPU holds for the life of the process, so the call makes no hipMalloc or
 * hipFree.  n is limited to 349,525 molecules (6 n doubles): larger systems
 * return G6_ECAPACITY.  The array is shared by every call on a GPU: calls
 * that overlap on one GPU race.  A tile system that runs out of a capacity
 * returns the error instead of falling back to the rows. */
int g6_eval_forces_noalloc(g6_handle h, double *force, double *torque, double *thermo);

/* ---- the reference's module entry points, one call each --------------- */
/* Grid_Update + Address_Assign + Grid_NeighborList + Interaction_Compress2
 * (GridModule.F90:95-125, 503-752; InteractionModule.F90:398-605), on the
 * COMs com[3n] of the handle's box.  Writes the compressed list in the
 * reference's order; *npair0 = candidates.  Returns G6_ECAPACITY if more than
 * cap pairs pass (nothing is written past cap). */
int g6_pairlist(g6_handle h, const double *com, int64_t cap, int64_t *npair0,
                int64_t *npair, int32_t *pair_i, int32_t *pair_j, double *ox,
                double *oy, double *oz);
/* Address_Assign cell of each molecule (0-based mAddr-1, mAddr = (z ny + y)
 * nx + x + 1) and the cells per axis of x (*ndiv; every axis of a cube). */
int g6_cells(g6_handle h, const double *com, int32_t *cell, int32_t *ndiv);
/* CutOff_Dimmer2 (CutoffModule.F90:220-278) for a given pair list. */
int g6_cutoff(g6_handle h, int64_t npair, const int32_t *pair_i,
              const int32_t *pair_j, const double *com, const double *ox,
              const double *oy, const double *oz, double *eratio,
              double *fratio);
/* force(iv,mi,mj,site,ep,vrmat,custom,si,sj) (Interaction2.F90:19-68):
 * site coordinates x,y,z[n*nsite] (site index i*nsite+s, CoM last);
 * site forces are ADDED to fx,fy,fz; ep[npair] and vrmat[9] (row-major
 * vrmat(a,b)) are overwritten.  Uses the handle's model and precision. */
int g6_force(g6_handle h, int64_t npair, const int32_t *pair_i,
             const int32_t *pair_j, const double *ox, const double *oy,
             const double *oz, const double *eratio, const double *fratio,
             const double *x, const double *y, const double *z, double *fx,
             double *fy, double *fz, double *ep, double *vrmat);
/* PairPotential / StdPairPotential energy-only path
 * (Interaction2.F90:71-121, StdInteraction.F90:633-762). */
int g6_pair_energy(g6_handle h, int64_t npair, const int32_t *pair_i,
                   const int32_t *pair_j, const double *ox, const double *oy,
                   const double *oz, const double *eratio, const double *x,
                   const double *y, const double *z, double *ep);
/* rigid_setsiteposition (RigidModule.F90:146-153): site coordinates from
 * com[3n] and quat1[4n] (quat(n,4) of order 1). */
int g6_site_positions(g6_handle h, const double *com, const double *quat1,
                      double *x, double *y, double *z);
/* rigid_collectforce (RigidModule.F90:275-317): molecule force[3n] and
 * body-frame torque[3n] from site forces and quat1[4n]. */
int g6_collect(g6_handle h, const double *quat1, const double *fx,
               const double *fy, const double *fz, double *force,
               double *torque);

/* ---- neighbour arrays of the order parameters (Neighbor.F90) ---------- */
/* orders of neighbors_module (Neighbor.F90:5) */
enum { G6_ORDER_NONE = 0, G6_ORDER_BY_DISTANCE = 1, G6_ORDER_BY_LABEL = 2 };
/* entries per array: MAXNEAR + 1 (Neighbor.F90:4, 8-13) */
#define G6_NEAR 151
/* A table of arrays is nnear[rows] and near, dx, dy, dz, dd[rows * G6_NEAR]
 * (Fortran near(G6_NEAR, rows)); near holds 0-based molecule indices and
 * (dx, dy, dz) the vector from the array's molecule to that neighbour.
 * radius <= 0 stands for the absent optional argument.  Each array is the
 * reference's: appended (ORDER_NONE) or sorted by distance or label with
 * ties in insertion order, cut to maxnei (0: no limit). */
/* neighbors_all2 (Neighbor.F90:91-159) over a pair list of nsite molecules
 * with centres coord[nsite*3] (Fortran coord(nsite,3)); nnear/near/dx/dy/
 * dz/dd have nsite rows.  G6_ECAPACITY when an array would need more than
 * G6_NEAR entries (the reference then writes past its arrays).  Uses the
 * handle's device and stream only. */
int g6_neighbors(g6_handle h, int nsite, int64_t npair, const int32_t *pair_i,
                 const int32_t *pair_j, const double *ox, const double *oy,
                 const double *oz, const double *coord, int maxnei, int order,
                 double radius, int32_t *nnear, int32_t *near, double *dx,
                 double *dy, double *dz, double *dd);
/* neighbor1_compress2 (Neighbor.F90:195-215) of each of nrow arrays: one
 * call replaces the reference's loop over molecules. */
int g6_neighbors_compress(g6_handle h, int nrow, const int32_t *nnear_in,
                          const int32_t *near_in, const double *dx_in,
                          const double *dy_in, const double *dz_in,
                          const double *dd_in, int maxnei, int order,
                          double radius, int32_t *nnear, int32_t *near,
                          double *dx, double *dy, double *dz, double *dd);
/* neighbors_displacement (Neighbor.F90:401-457): molecule `target` of a
 * table of nsite arrays built with `order` moves by (ddx, ddy, ddz); writes
 * the *nmodified updated arrays (the moved molecule's first, then those of
 * its neighbours that list it, in its array's order) and their molecule
 * indices index[]; the output table needs nnear_in[target] + 1 <= 152 rows. */
int g6_neighbors_displacement(g6_handle h, int nsite, const int32_t *nnear_in,
                              const int32_t *near_in, const double *dx_in,
                              const double *dy_in, const double *dz_in,
                              const double *dd_in, int order, int target,
                              double ddx, double ddy, double ddz,
                              int32_t *nmodified, int32_t *index, int32_t *nnear,
                              int32_t *near, double *dx, double *dy, double *dz,
                              double *dd);

/* ---- umbrella order parameters of mc2 (Q6Local.F90, Triangle.F90) -------
 * On a pair list and molecule centres as g6_neighbors.  Q6 local bond order,
 * sq6system_prepareall (Q6Local.F90:101-231): q6[nsite], the bond count
 * nbond[nsite] of each molecule, their sum and op = sum / nsite;
 * G6_ECAPACITY when a molecule has more than MAXBOND = 10 neighbours within
 * 3 A (the reference's arrays end there). */
int g6_q6local(g6_handle h, int nsite, int64_t npair, const int32_t *pair_i,
               const int32_t *pair_j, const double *ox, const double *oy,
               const double *oz, const double *coord, double *q6, int32_t *nbond,
               double *sumq6, double *op);
/* sTriangle_prepareall (Triangle.F90:53-116): the histogram (21 bins of 0.1
 * from -1) of the O-O-O angles of each molecule's four nearest neighbours
 * within 5 A, and its Kullback-Leibler divergence from refhist[21]
 * (DistribModule.F90:304-326); logarithmic: its log (LOGTRIAN). */
int g6_triangle(g6_handle h, int nsite, int64_t npair, const int32_t *pair_i,
                const int32_t *pair_j, const double *ox, const double *oy,
                const double *oz, const double *coord, const double *refhist,
                int logarithmic, double *hist, double *op);
/* The order parameters after a trial displacement of molecule target
 * (0-based) by (dx, dy, dz), each component below 1 A, on the same pair list
 * and image shifts: what sq6system_Differentiate (Q6Local.F90:236-465) and
 * sTriangle_Differentiate (Triangle.F90:125-249) return to mc2's umbrella
 * step (mc2.F90:690-692), evaluated in full on the displaced centres.  The
 * reference updates its sums incrementally and agrees with this except when
 * two molecules bonded to each other (within 3 A) both enter or leave the
 * target's 5 A shell in one move: its Q6 update then adds their bond to the
 * target's sum from both sides, and this returns its own full evaluation
 * (sq6system_prepareall) of the displaced configuration instead. */
int g6_q6local_displacement(g6_handle h, int nsite, int64_t npair,
                            const int32_t *pair_i, const int32_t *pair_j,
                            const double *ox, const double *oy, const double *oz,
                            const double *coord, int target, double dx, double dy,
                            double dz, double *q6, int32_t *nbond, double *sumq6,
                            double *op);
int g6_triangle_displacement(g6_handle h, int nsite, int64_t npair,
                             const int32_t *pair_i, const int32_t *pair_j,
                             const double *ox, const double *oy, const double *oz,
                             const double *coord, int target, double dx, double dy,
                             double dz, const double *refhist, int logarithmic,
                             double *hist, double *op);

/* ---- constant-pressure MD: the Andersen barostat (AndersenModule.F90) ---
 * mode 1 (orthorhombic: every edge scales by (v0/V)^(1/3)) or mode 2 (orz:
 * the z edge alone takes the volume, and the dilation drag of the corrector
 * acts on z alone, RigidModule.F90:1290-1306) with external pressure pex
 * (atm) and mass per molecule as @PCN4 (Mixture.F90:1207-1231); mode 0
 * switches it off.
 * v: NULL starts the barostat (Andersen_Initialize: v0 = the box volume),
 * or v[0..5] = v0..v5 continue a saved state (@PCN5).  g6_step then runs
 * Andersen_Predict before the forces and Andersen_Correct with the step's
 * pressure after the corrector (Mixture.F90:1843-1844, 2283-2286); the box
 * edge changes every step.  The thermo record adds the barostat energies
 * ([30], [31]) to H ([13]); [25] is the pressure in the step's box, [26]
 * the corrected volume.  The module entry points (g6_pairlist, g6_force,
 * ...) keep the box of g6_init.  Not for handles in a ladder. */
int g6_set_andersen(g6_handle h, int mode, double pex, double mass,
                    const double *v);
/* pex, mass (per molecule, as @PCN4), v0..v5: the @PCN5 payload */
int g6_get_andersen(g6_handle h, double *state8);
/* the current box edge (A; the x edge of an orthorhombic box) */
int g6_get_box(g6_handle h, double *box_l);
/* the current box edges edges[3] (A) */
int g6_get_box3(g6_handle h, double *edges);

/* ---- Monte Carlo: PairInfo and the mc2 trial loop ------------------- */
/* Starts an MC run on the handle's current configuration (COMs and order-1
 * quaternions of the last g6_upload_state): site positions, the PairInfo
 * matrix of every pair within rc (pairenergy, PairInfo.F90:12-46, built as
 * mc2.F90:303-340 builds it, StdPairPotential energies times S), its sum,
 * per-molecule step sizes delta[n] and rotation ratios rtratio[n] (@DELT),
 * beta = 1/temp (@TEMP) and the random stream seeded with seed (@SEED).
 * The models whose pair energy is StdPairPotential's (TIP4P, TIP5P, ST2 as
 * TIP5P, SPC/E, NvdE, OPLS methanol); n <= 1500 (the trial loop keeps the
 * configuration in LDS).  Clears any umbrellas. */
int g6_mc_begin(g6_handle h, double temp, uint32_t seed, const double *delta,
                const double *rtratio);
/* ntrial single-molecule trials of mc2 (mc2.F90:540-862): target from the
 * stream, Gaussian rotation (qinfrotator + qadd) and displacement with
 * widths delta*rtratio and delta, dE of the target against every molecule
 * within rc + delta, Metropolis at beta, PairInfo and step-size updates.
 * *naccept = accepted trials since g6_mc_begin. */
int g6_mc_trials(g6_handle h, int ntrial, long long *naccept);
/* The MC configuration: com[3n], quat1[4n], delta[n], rtratio[n] and the
 * running energy sum (any pointer may be NULL). */
int g6_mc_get(g6_handle h, double *com, double *quat1, double *delta,
              double *rtratio, double *epsum);
/* The PairInfo matrix ep[n*n] (ep[i*n+j], symmetric) and the running sum. */
int g6_mc_pairinfo(g6_handle h, double *ep, double *epsum);
/* The total energy of the current MC configuration from a fresh PairInfo
 * (the bookkeeping is left as it is). */
int g6_mc_energy(g6_handle h, double *epsum);

/* Umbrella sampling in the trial loop: mc2 built with OP (mc2.F90:18).
 * Kinds of @GEUM (mc2util.F90:174-191): the triangle order parameter
 * (sTriangle, Triangle.F90:53-280): the Kullback-Leibler divergence
 * (DistribModule.F90:304-326) from refhist[21] of the histogram of O-O-O
 * angle cosines (21 bins of 0.1 from -1) between each molecule's four
 * nearest neighbours among those within 5 A at the loop's start, or its
 * log (LOGTRIAN); the local bond order Q6 (Q6LOCAL_, sq6system,
 * Q6Local.F90:42-750): the mean over molecules of q6 from the bonds within
 * 3 A inside each molecule's 5 A shell, its sums updated over the trials of
 * a loop as the reference updates them (rejected trials included, so the
 * value is the reference's to the last bits of its sums); the energy
 * (EPOP____). */
enum { G6_UMB_TRIANGLE = 1, G6_UMB_Q6LOCAL = 2, G6_UMB_EPOP = 3, G6_UMB_LOGTRIAN = 4 };
/* numbrella (<= 4; 0 removes them) potentials coeffi (op - center)^2
 * (Umbrella.F90:11-16), refhist[21] for the triangle kinds (@HIST, read
 * as mc2util.F90:156-159).  With umbrellas set, every g6_mc_trials call is
 * one mc2 loop: it first prepares the order parameters on the current
 * configuration (the 5 A tables from the pairs within rc, mc2.F90:486-521;
 * op = the energy for EPOP____), then each trial adds
 * sum(U(newop) - U(op)) to dE before the Metropolis test (mc2.F90:684-706)
 * and an accepted trial moves the tables (sTriangle_Integrate,
 * sq6system_Integrate / _Reject).  A molecule with more than 64 molecules
 * within 5 A (triangle) or 6 A (Q6), or with more than MAXBOND = 10 bonds,
 * stops the call before any trial: G6_ECAPACITY, the state unchanged.  A
 * trial that would give a molecule an eleventh bond (the reference's arrays
 * end there) stops the call before that trial: G6_ECAPACITY, the trials
 * before it kept.  Requires g6_mc_begin. */
int g6_mc_set_umbrella(g6_handle h, int numbrella, const int32_t *kind, const double *center,
                       const double *coeffi, const double *refhist);
/* op[numbrella]: each order parameter after the last g6_mc_trials call
 * (what mc2's @MC2O line prints, mc2.F90:876-884; 0 before the first);
 * center and coeffi as set (any pointer may be NULL). */
int g6_mc_get_umbrella(g6_handle h, double *op, double *center, double *coeffi);
/* beta = 1/temp for the following trials (an exchange's new temperature) */
int g6_mc_set_temperature(g6_handle h, double temp);

/* ---- mixtures of two rigid components (@NCMP 2 and two @ID08/@NX4A
 * blocks in a Mixture deck; Mixture.F90:636-712 reads them, 1833-2290 steps
 * them).  The components are two handles of g6_init / g6_init_box3 with one
 * box, rc, taper, dt, skin and precision, neither on the barostat; a
 * component on the cell tiles moves to the rows.  The pairs inside each
 * component keep that component's
 * force (g6_set_fast); the pairs between them are StdForce's
 * (StdInteraction.F90:63-630, Interaction2.F90:19-67 without @FAST) with
 * the Lorentz-Berthelot site parameters of the two models, every pair
 * whose COM distance is within rc, switched on the COM distance, in FP64
 * from the site offsets of the records (FP32 ones under G6_FP32).  One Nose
 * thermostat (the first component's, g6_set_nose) over 6 (nA + nB) degrees
 * of freedom drives both.  A component takes part in one mixture at a time
 * and should not be stepped on its own while in it. */
typedef int64_t g6_mixture;

int g6_mixture_create(g6_mixture *hm, g6_handle first, g6_handle second);
int g6_mixture_free(g6_mixture hm);
/* Rigid_VReset2 of the first component, then of the second, from one
 * random stream (Mixture.F90:1336-1346) */
int g6_mixture_vreset(g6_mixture hm, double temp, uint32_t seed);
/* nsteps MD steps of both components.  G6_ECAPACITY if a capacity of
 * either component ran out (its state is then not the reference's). */
int g6_mixture_step(g6_mixture hm, int nsteps);
/* thermo record of the whole system (layout of g6_get_thermo; thermo[0]
 * the total pair energy) and, if parts is non-null, parts[5]: the pair
 * energy inside the first component, inside the second, between them, and
 * the kinetic energy of each component (internal units) -- the ep(k) and
 * ekt + ekr columns of Mixture's @LOGD line (Mixture.F90:2453-2460). */
int g6_mixture_thermo(g6_mixture hm, double *thermo, double *parts);

/* ---- replica ladders (MDReplica.F90, the replica branch of Mixture.F90)
 * The reference runs one process per replica (fakempi) and exchanges
 * temperatures every @INNR steps.  Here the replicas are systems of one
 * process: a ladder steps all of them together, each MD kernel launched
 * once for every system, and an exchange group applies the reference's
 * swap rule between them. */
typedef int64_t g6_ladder;
typedef int64_t g6_rx;

/* A ladder over nsys distinct systems (g6_init handles) of one model and
 * precision with equal rc and taper.  The systems keep their own state and
 * stay usable on their own between ladder calls (on the rows of both
 * directions from here on). */
int g6_ladder_create(g6_ladder *hl, int nsys, const g6_handle *hs);
int g6_ladder_free(g6_ladder hl);
/* nsteps MD steps of every system: what g6_step does to each, with the
 * same results bit for bit (the inner loop of Mixture.F90:1833-2286 on
 * every node), launched as one captured hipGraph. */
int g6_ladder_step(g6_ladder hl, int nsteps);
/* Block assignment of the ladder's launches (the autotuner's knob; results
 * do not depend on it): G6_LADDER_MAP_FLAT lays the systems' blocks one
 * after another (the default); G6_LADDER_MAP_XCD gives every block of a
 * system the same XCD (workgroup b runs on XCD b mod 8; systems balanced
 * over the eight, largest first). */
enum { G6_LADDER_MAP_XCD = 0, G6_LADDER_MAP_FLAT = 1 };
int g6_ladder_set_map(g6_ladder hl, int mode);
/* The ladder's force kernel alone (one launch covers every system), nlaunch
 * launches back to back on the state a g6_ladder_step call leaves, timed
 * with HIP events on the ladder's stream; *ms = mean milliseconds per
 * launch.  The state is not moved.  No reference counterpart: the
 * measurement hook of the ladder's roofline. */
int g6_ladder_time_force(g6_ladder hl, int nlaunch, double *ms);
/* the thermo records of the last step: thermo[nsys * G6_THERMO_LEN], in
 * the order of hs */
int g6_ladder_thermo(g6_ladder hl, double *thermo);

/* Thermostat retarget after an exchange (Mixture.F90:1630-1653 with
 * Rigid_ScaleVelocity, RigidModule.F90:621-633): dx(1) and w(1) scaled by
 * sqrt(temp / nose%temp), nose%temp = temp, zeta0..zeta4 = 0.  Requires an
 * active Nose thermostat. */
int g6_set_temperature(g6_handle h, double temp);

/* An exchange group of nrep replicas (node k = hs[k-1]; hs may be NULL for
 * the host rule alone): directed bonds a[k] -> b[k], 1-based, in the order
 * graph2pair gives (Graph.F90:34-52; @RXPR chain: 1->2, 2->1, 2->3, ...);
 * nexchange trials per exchange (@NRXG); MT19937 seed of the group's
 * stream (@SEED).  node_in_charge starts as the identity (Mixture.F90:1483). */
int g6_rx_create(g6_rx *hx, int nrep, const g6_handle *hs, int nbond, const int32_t *a,
                 const int32_t *b, int nexchange, uint32_t seed);
int g6_rx_free(g6_rx hx);
/* The exchange loop of Mixture.F90:1605-1622 with replica_exchange_trial
 * (Replica.F90:7-73, no umbrella): energy[nrep] (internal units) and
 * beta[nrep] (1/K) indexed by node; beta is updated in place and
 * node_in_charge kept by the group.  naccept = accepted swaps. */
int g6_rx_trials(g6_rx hx, const double *energy, double *beta, int32_t *naccept);
/* The same trials weighted by umbrellas (Replica.F90:30-49), as mc2's
 * replica exchange of MC nodes runs them (mc2.F90:973-996): op, center and
 * coeffi hold numbrella entries per node, node k at (k-1)*numbrella.  A
 * state's weight adds coeffi (op - center)^2 of each umbrella to the
 * node's energy; the swapped state evaluates each node's order parameters
 * under the other node's umbrellas; an accepted swap exchanges the two
 * nodes' centers and coefficients (updated in place) with their betas.
 * The arguments mean what Replica.F90 names them; mc2.F90:981 passes its
 * beta array in the energy position and its energies in the beta position,
 * and a caller reproducing mc2's exchanges passes them the same way. */
int g6_rx_trials_umbrella(g6_rx hx, const double *energy, double *beta, int numbrella,
                          const double *op, double *center, double *coeffi,
                          int32_t *naccept);
/* node_in_charge[nrep] (1-based, as the reference) */
int g6_rx_nodes(g6_rx hx, int32_t *node_in_charge);
/* One exchange of Mixture.F90:1570-1653 between the group's systems:
 * energies thermo[0] and temperatures thermo[29] of each replica's record
 * (thermo[nrep * G6_THERMO_LEN] in node order, e.g. from g6_ladder_thermo;
 * NULL reads each system's own), beta = 1/temp, the trials, then
 * g6_set_temperature(1/beta) on every replica whose beta changed. */
int g6_rx_exchange(g6_rx hx, const double *thermo, int32_t *naccept);

/* ---- NPT Monte Carlo with collective moves (no reference counterpart: the
 * reference's MC has single-molecule NVT moves; the energy is PairInfo,
 * PairInfo.F90:12-46, as g6_mc_pairinfo) ------------------------------- */
/* Starts an NPT run on the handle's current configuration (COMs and order-1
 * quaternions of the MD state, box L of g6_init): temp (K), press_bar,
 * seed of the proposal stream (Philox4x32-10), block moves of every molecule
 * of one cell of the Address_Assign grid (ndiv = int(L/voxa)) shifted by a
 * random offset per move, displaced within a cube of half-diagonal disp_max
 * (A) and turned by at most rot_max_deg, volume moves ln V +- dlnv_max
 * (<= 0.03) with probability p_volume.  A block move that takes a member out
 * of the shifted cell, a block of more than 96 molecules and a volume move
 * to a box edge below 2 rc are void: rejected untested (the reverse of a
 * void move is void, so detailed balance holds).  n <= 4096, TIP4P or TIP5P,
 * rc < L (1 - dlnv_max) / 2 at the start. */
int g6_npt_begin(g6_handle h, double temp, double press_bar, uint64_t seed, double disp_max,
                 double rot_max_deg, double dlnv_max, double p_volume, double voxa);
/* hm = 0: count plain moves, each decided on the exact FP64 energy.
 * hm = a g6_mlp: count delayed-acceptance rounds of K (1..64) proposals
 * screened by the network (per-molecule energies of 32 features within
 * 6 A), the first that passes decided on the exact energy.  The surrogate
 * is a function of the configuration only while no O-O pair can come within
 * 6 A from beyond the round's 6.5 A list: G6_EARG if twice the largest O
 * move of a block (disp_max plus the turn of the O about the COM) reaches
 * 0.5 A.
 * stats[10] (cumulative): proposals, accepted, volume proposals, volume
 * accepted, surrogate passes, exact evaluations, rounds, void proposals,
 * next proposal index, volume moves among the surrogate passes.  G6_ECAPACITY if a molecule had more than 96
 * neighbours within 6.5 A for the descriptors. */
int g6_npt_run(g6_handle h, int count, int64_t hm, int K, long long *stats);
/* com[3n], quat1[4n], box edge, PairInfo energy sum (internal units) */
int g6_npt_get(g6_handle h, double *com, double *quat1, double *box_l, double *energy);
/* the PairInfo sum recomputed from scratch on the current configuration */
int g6_npt_energy(g6_handle h, double *energy);
/* features x[n][32] (float) of the current configuration and the exact
 * per-molecule energies e[n] = (1/2) sum_j pinfo(i, j) */
int g6_npt_features(g6_handle h, float *x, double *e);

/* ---- learned energy surrogate (no reference counterpart) ---------------
 * A per-molecule MLP 32 -> 64 -> 64 -> 1 with SiLU: BF16 weights on the
 * matrix cores, FP32 accumulation, FP32 master weights.  params[6337] =
 * W1[64][32] (output-major), b1[64], W2[64][64], b2[64], w3[64], b3;
 * norm[66] = feature mean[32], feature 1/std[32], y_shift, y_scale: the
 * network sees (x - mean) / std and its output o gives y_scale*o + y_shift. */
typedef int64_t g6_mlp;
#define G6_MLP_NIN 32
#define G6_MLP_NPARAM 6337
#define G6_MLP_NNORM 66
/* zero weights, identity normalisation */
int g6_mlp_create(g6_mlp *hm);
int g6_mlp_free(g6_mlp hm);
/* either pointer may be NULL (left unchanged); new params restart Adam */
int g6_mlp_set_params(g6_mlp hm, const float *params, const float *norm);
int g6_mlp_get_params(g6_mlp hm, float *params, float *norm);
/* y[m] for feature rows x[m][32] */
int g6_mlp_forward(g6_mlp hm, int64_t m, const float *x, float *y);
/* loss = mean over rows of (o - (t - y_shift)/y_scale)^2 and its gradient
 * grad[6337] with respect to params */
int g6_mlp_loss_grad(g6_mlp hm, int64_t m, const float *x, const float *t, float *loss,
                     float *grad);
/* Adam (0.9, 0.999, 1e-8) on minibatches of a shuffled copy of the rows
 * (MT19937 seed), data resident on the device; loss_per_epoch[epochs] */
int g6_mlp_train(g6_mlp hm, int64_t m, const float *x, const float *t, int epochs, int batch,
                 float lr, uint32_t seed, float *loss_per_epoch);

/* ---- synthetic start (no reference counterpart: the reference reads
 * @NX4A files) --------------------------------------------------------- */
/* Simple-cubic lattice, m = ceil(n^(1/3)) per edge, occupied sites chosen by
 * an MT19937-seeded shuffle, Shoemake-uniform quaternions.  com[3n], quat1[4n].*/
int g6_lattice_start(int n, double box_l, uint32_t seed, double *com,
                     double *quat1);

/* ---- info -------------------------------------------------------------- */
/* nsteps MD steps launched one kernel at a time (no graph) with HIP events
 * on the library's stream around each phase; t[0..3] = mean ms per step of
 * the force kernel, the pair-list build, the per-molecule kernels and the
 * whole step; t[4] = mean interacting pairs per step.  Same results as
 * g6_step. */
int g6_step_timed(g6_handle h, int nsteps, double *t);
/* The force kernel alone: nlaunch launches back to back in one hipGraph
 * (eager launches when G6_GRAPH_STEPS=0) on the state a g6_step call leaves
 * (same forces each time), timed with HIP events on the library stream;
 * *ms = mean milliseconds per launch.  The state is not moved.  Not with
 * the barostat. */
int g6_time_force(g6_handle h, int nlaunch, double *ms);
/* The force kernel a system's MD path runs (no counterpart in the reference,
 * whose force routine has one form): info[0] = 1 for the cell-tile kernel
 * that evaluates each pair once, 0 for the rows of both directions;
 * info[1] = tile cells per axis, info[2] = workgroups per tile, info[3] =
 * waves per workgroup, info[4] = neighbourhood slots per tile (zeros for the
 * rows), info[5] = the capacity code that moved a tile system to the rows
 * (3 cell residents, 4 neighbourhood slots or row entries, 2 image shift, 6
 * a pair force beyond the fixed-point sums' range; 0 when none did).  FP32 fixed-box systems with at least 256 cells of edge
 * rc + skin take the tiles; G6_TILES=0 in the environment at g6_init keeps
 * the rows and G6_TILE_MIN sets the cell threshold.  A system that a
 * capacity error moved to the rows tries the tiles again at the first
 * g6_step call 1000 steps later (G6_TILE_RETRY), the gap doubling after
 * each failed try; systems in a ladder or under the barostat keep the rows. */
int g6_md_kernel(g6_handle h, int32_t *info);
int g6_model_nsite(int model);

#ifdef __cplusplus
}
#endif
#endif
