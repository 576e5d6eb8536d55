/*
 * wadges.h -- C ABI of the MI355X (gfx950) WADGES right-hand-side / time-step
 * hot path: hand-written HIP kernels for the volume flux-differencing, surface
 * entropy-stable flux and weight-adjusted inverse-mass + LSRK update kernels.
 *
 * Drop-in boundary.  The reference driver launches three OKL kernels per RK
 * stage through OCCA (euler.c:3610-3670):
 *
 *   vol    (K, vgeo, vfgeo, nrJ, nsJ, [ntJ], Drq, Dsq, [Dtq, Drstq], VqLq, VfPq,
 *           Q, Qf, rhsQ, rhsQf)                     euler.c:3615-3617 / :3656-3658
 *   surf   (K, fgeo, mapPq, VqLq, Qf, rhsQf, rhsQ)  euler.c:3619-3620 / :3660-3661
 *   update (K, Jq, VqPq, VfPq, fa, fb, fdt, rhsQ, resQ, Q, Qf)
 *                                                   euler.c:3622-3625 / :3663-3666
 *
 * wadges_vol / wadges_surf / wadges_update keep those roles, argument order and
 * array layouts; the OCCA compile-time defines (p_Nq, p_Nfq, p_Nfaces,
 * p_Nfields, p_gamma, p_TAU, p_Kblk*; euler.c:2880-2948) become a plan.  Every
 * pointer is a device pointer, every call is asynchronous on the given HIP
 * stream (a hipStream_t passed as void*, NULL = default stream), no call
 * allocates, and the calls can be captured in a HIP graph.
 *
 * The volume-type calls (wadges_vol, wadges_stage, wadges_rhs_vjp) read the
 * operators through tables the plan re-lays out on the device.  The tables are
 * prepared on the first such call, and again whenever the operator pointers
 * (Drq, Dsq, Dtq, VqLq, VfPq) differ from those of the previous preparation;
 * that call waits for the preparation.  Captured calls (HIP stream capture)
 * read a second set of tables, whose preparation is recorded into the graph
 * at the capture's first call with those operators: a graph rewrites the
 * tables it reads whenever it runs and never touches those of the calls
 * outside capture.  Graphs of one plan therefore must not run concurrently
 * with each other.  After changing operator values in place, call
 * wadges_plan_invalidate.
 *
 * Layouts (Appendix of the reference's operators_c.cpp:322-389 and
 * euler.c:3158-3170), all FP64 except mapPq:
 *   Q, rhsQ, resQ [K][Nfields][Nq]     Qf, rhsQf [K][Nfields][Nf], Nf = Nfq*Nfaces
 *   vgeo [K][Nvgeo][Nq]  vfgeo [K][Nvgeo][Nf]  fgeo [K][Nfgeo][Nf]  Jq [K][Nq]
 *   mapPq [K][Nf] (uint32): e_nbr*Nf*Nfields + neighbour face point
 *   Drq, Dsq, Dtq, VqPq [Nq x Nq], VqLq [Nq x Nf], VfPq [Nf x Nq], column major
 *   nrJ, nsJ, ntJ [Nf]
 * with Nfields = dim+2, Nvgeo = dim*dim, Nfgeo = dim+1, Nfaces = dim+1.
 *
 * Return codes: 0 on success, negative on error (see WADGES_E*); the library
 * never aborts the process.
 */
#ifndef WADGES_H
#define WADGES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  WADGES_OK = 0,
  WADGES_EINVAL = -1,       /* bad argument (null pointer, K < 0, size mismatch) */
  WADGES_EUNSUPPORTED = -2, /* (dim, Nq, Nfq) has no compiled kernel instance   */
  WADGES_EHIP = -3,         /* a HIP runtime call failed (wadges_last_hip_error) */
  WADGES_ENOMEM = -4,       /* device allocation at plan creation failed         */
  WADGES_ETOOLARGE = -5     /* size beyond a fixed capacity (wadges_diagnostics_noalloc) */
};

typedef struct wadges_plan wadges_plan;

/* Creates a plan: the reference's kernel defines (euler.c:2880-2948) as a
 * runtime object.  dim = 2 (triangles) or 3 (tetrahedra); N the polynomial
 * degree; Nq, Nfq the volume / per-face quadrature sizes the operators were
 * built with; gamma = p_gamma; tau = p_TAU (the reference's 2D surface kernel
 * has no tau and equals tau = 1).  Allocates the plan's small device workspace
 * on the current device; of the single-problem calls this is the only one that
 * allocates (the ensemble calls wadges_batch_create, _set_config and _tune
 * allocate too). */
int wadges_plan_create(wadges_plan **plan, int dim, int N, int Nq, int Nfq, double gamma,
                       double tau);
int wadges_plan_destroy(wadges_plan *plan);
/* Forgets the prepared operator tables (see above); the next call prepares them. */
int wadges_plan_invalidate(wadges_plan *plan);

/* Volume kernel: role of euler_vol_3d_curved (okl/Euler3D.okl:2123-2523),
 * euler_vol_2d_curved (okl/Euler2D.okl:773-1353) and, on affine meshes, of
 * euler_vol_3d (okl/Euler3Daffine.okl:152, same arguments).  Reads the entropy-projected
 * volume state from rhsQ and the face state from Qf; overwrites rhsQ with the
 * flux-differencing volume term and writes the volume-to-face coupling term to
 * rhsQf.  Q and Drstq are accepted for argument compatibility and not read (the
 * reference kernel does not read them either).  2D: ntJ, Dtq, Drstq may be NULL. */
int wadges_vol(const wadges_plan *plan, int K, const double *vgeo, const double *vfgeo,
               const double *nrJ, const double *nsJ, const double *ntJ, const double *Drq,
               const double *Dsq, const double *Dtq, const double *Drstq, const double *VqLq,
               const double *VfPq, const double *Q, const double *Qf, double *rhsQ,
               double *rhsQf, void *stream);

/* Surface kernel: role of euler_surf_3d (okl/Euler3D.okl:3242-3473; the
 * affine euler_surf_3d, okl/Euler3Daffine.okl:421, has the same arguments) and
 * euler_surf_2d (okl/Euler2D.okl:1357-1495).  Entropy-conservative flux plus
 * Lax-Friedrichs penalty at every face point, plus rhsQf, lifted by VqLq and
 * added to rhsQ. */
int wadges_surf(const wadges_plan *plan, int K, const double *fgeo, const uint32_t *mapPq,
                const double *VqLq, const double *Qf, const double *rhsQf, double *rhsQ,
                void *stream);

/* Update kernel: role of euler_update_3d_curved (okl/Euler3D.okl:2527-2884),
 * euler_update_2d_curved (okl/Euler2D.okl:1729-2282) and, on affine meshes, of
 * euler_update_3d (okl/Euler3Daffine.okl:555, same arguments).  Weight-adjusted inverse
 * mass -2 VqPq diag(1/J) VqPq, low-storage RK update of resQ and Q, then the
 * weight-adjusted entropy projection: Qf = u(VfPq v~), rhsQ = u(VqPq v~). */
int wadges_update(const wadges_plan *plan, int K, const double *Jq, const double *VqPq,
                  const double *VfPq, double fa, double fb, double fdt, double *rhsQ,
                  double *resQ, double *Q, double *Qf, void *stream);

/* One whole LSRK stage in a single launch: vol + surf + update fused per
 * element (no counterpart kernel in the reference; it replaces the three
 * launches of rk_step, euler.c:3656-3666, with one).  The volume RHS stays on
 * chip.  Because neighbouring elements read face traces of the previous stage
 * while an element writes its new ones, the face traces are double-buffered:
 * the stage reads Qf_in and writes Qf_out (must differ; swap them between
 * stages).  rhsQ (projected volume state in and out), resQ and Q are updated
 * in place as by wadges_update.  Results equal vol -> surf -> update to
 * round-off. */
int wadges_stage(const wadges_plan *plan, int K, const double *vgeo, const double *vfgeo,
                 const double *fgeo, const uint32_t *mapPq, const double *nrJ, const double *nsJ,
                 const double *ntJ, const double *Drq, const double *Dsq, const double *Dtq,
                 const double *VqLq, const double *VqPq, const double *VfPq, const double *Jq,
                 double fa, double fb, double fdt, const double *Qf_in, double *Qf_out,
                 double *rhsQ, double *resQ, double *Q, void *stream);

/* wadges_stage with separate input and output arrays for the projected state,
 * the RK residual and the state (each output may equal its input; Qf_out must
 * differ from Qf_in).  Same arithmetic as wadges_stage; it lets a caller keep
 * every stage of a trajectory without copies (the adjoint's recomputation).
 * No counterpart in the reference. */
int wadges_stage_record(const wadges_plan *plan, int K, const double *vgeo, const double *vfgeo,
                        const double *fgeo, const uint32_t *mapPq, const double *nrJ,
                        const double *nsJ, const double *ntJ, const double *Drq, const double *Dsq,
                        const double *Dtq, const double *VqLq, const double *VqPq,
                        const double *VfPq, const double *Jq, double fa, double fb, double fdt,
                        const double *Qf_in, double *Qf_out, const double *rhsQ_in,
                        double *rhsQ_out, const double *resQ_in, double *resQ_out,
                        const double *Q_in, double *Q_out, void *stream);

/* ---- Discrete adjoint (vector-Jacobian products); no counterpart in the
 * reference.  Cotangent arrays use the layout of their primal arrays.
 *
 * wadges_stage_vjp: VJP of one whole stage (wadges_stage or wadges_stage_record
 * with coefficients fa, fb, fdt) in two launches: wadges_update_vjp with the
 * transposed lift s_bar = VqLq^T R_bar fused in, then the RHS VJP.  rhsQ_in,
 * Qf_in are the stage's input projected state, Q_out its output state.  On
 * entry Q_bar, resQ_bar, rhsQ_bar_out, Qf_bar_out are the cotangents of the
 * stage outputs; on return Q_bar, resQ_bar are those of the stage inputs Q,
 * resQ and rhsQ_bar_in, Qf_bar_in those of rhsQ_in, Qf_in.  R_bar (K*Nfields*Nq)
 * and sbar (K*Nfields*Nf) are caller scratch.  rhsQ_bar_in may equal
 * rhsQ_bar_out and Qf_bar_in may equal Qf_bar_out (the second launch writes
 * them after the first has read them). */
int wadges_stage_vjp(const wadges_plan *plan, int K, const double *vgeo, const double *vfgeo,
                     const double *fgeo, const uint32_t *mapPq, const double *nrJ,
                     const double *nsJ, const double *ntJ, const double *Drq, const double *Dsq,
                     const double *Dtq, const double *VqLq, const double *VqPq, const double *VfPq,
                     const double *Jq, double fa, double fb, double fdt, const double *rhsQ_in,
                     const double *Qf_in, const double *Q_out, double *Q_bar, double *resQ_bar,
                     const double *rhsQ_bar_out, const double *Qf_bar_out, double *R_bar,
                     double *sbar, double *rhsQ_bar_in, double *Qf_bar_in, void *stream);

/* The single-kernel VJPs the stage VJP is made of:
 *
 * wadges_rhs_vjp: given rhsQ_bar, the cotangent of the stage RHS (the rhsQ that
 * vol -> surf produce from the projected state rhsQ and traces Qf), writes the
 * cotangents of that state: rhsQ_state_bar [K][Nfields][Nq] and Qf_bar
 * [K][Nfields][Nf].  sbar is caller scratch of K*Nfields*Nf doubles.  The log
 * mean is differentiated in the branch the forward takes; max(lamM, lamP) of
 * the Lax-Friedrichs penalty passes its derivative to the larger argument (to
 * the element's own point at ties). */
int wadges_rhs_vjp(const wadges_plan *plan, int K, const double *vgeo, const double *vfgeo,
                   const double *fgeo, const uint32_t *mapPq, const double *nrJ, const double *nsJ,
                   const double *ntJ, const double *Drq, const double *Dsq, const double *Dtq,
                   const double *VqLq, const double *VfPq, const double *rhsQ, const double *Qf,
                   const double *rhsQ_bar, double *sbar, double *rhsQ_state_bar,
                   double *Qf_bar, void *stream);

/* wadges_update_vjp: VJP of wadges_update.  Q is the state AFTER the update.
 * On entry Q_bar, resQ_bar hold the cotangents of the updated Q, resQ and
 * rhsQ_bar_out, Qf_bar_out those of the updated projected state; on return
 * Q_bar, resQ_bar hold the cotangents of Q, resQ before the update and rhsQ_bar
 * that of the RHS the update consumed. */
int wadges_update_vjp(const wadges_plan *plan, int K, const double *Jq, const double *VqPq,
                      const double *VfPq, double fa, double fb, double fdt, const double *Q,
                      double *Q_bar, double *resQ_bar, const double *rhsQ_bar_out,
                      const double *Qf_bar_out, double *rhsQ_bar, void *stream);

/* wadges_project_vjp: Q_bar += VJP of the initial projection
 * rhsQ = u(VqPq v(Q)), Qf = u(VfPq v(Q)) (euler.c:4028-4133) at Q. */
int wadges_project_vjp(const wadges_plan *plan, int K, const double *VqPq, const double *VfPq,
                       const double *Q, const double *rhsQ_bar, const double *Qf_bar,
                       double *Q_bar, void *stream);

/* Diagnostics as device reductions (the reference computes them on the host or
 * in euler_3d_compute_aux, okl/Euler3D.okl:3888-3934, euler.c:4135-4156 and
 * :4227-4274).  Writes per element aux[K][4] = (kinetic energy
 * sum wq J |m|^2/(2 rho), entropy sum wq J (-rho s), squared L2 error
 * sum wq J |Q - Qexact|^2 (0 if Qexact is NULL), mass sum wq J rho), and their
 * sums over the elements to totals[4], in a fixed summation order.  aux is
 * caller-provided scratch of K*4 doubles. */
int wadges_diagnostics(const wadges_plan *plan, int K, const double *wq, const double *Jq,
                       const double *Q, const double *Qexact, double *aux, double *totals,
                       void *stream);

/* wadges_diagnostics without the aux argument: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-element partials live in a fixed device array of the library,
 * WADGES_NOALLOC_MAX_K * 4 doubles (8 MiB) per GPU, held until the process
 * ends; its address is looked up once per device at plan creation.  K above
 * WADGES_NOALLOC_MAX_K returns WADGES_ETOOLARGE.  Same kernels and summation
 * order as wadges_diagnostics, so the totals are bit for bit the same.  The
 * call allocates nothing.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (different streams, or different host threads) race and
 * must be ordered by the caller. */
#define WADGES_NOALLOC_MAX_K 262144
int wadges_diagnostics_noalloc(const wadges_plan *plan, int K, const double *wq,
                               const double *Jq, const double *Q, const double *Qexact,
                               double *totals, void *stream);

/* ---- Ensembles: many independent meshes (uneven K and N, one dim) advanced
 * together, one launch per LSRK stage for the whole ensemble.  No counterpart
 * in the reference, which runs one mesh per process through the same stage
 * sequence (rk_run, euler.c:3672-3699); per problem the arithmetic is exactly
 * that of wadges_stage.
 *
 * Each problem is described by device pointers in the layouts above.  Its face
 * traces are double-buffered: Qf[0] holds the current traces at create, and
 * wadges_batch_current tells which of Qf[0], Qf[1] holds them after any number
 * of stages.  Operator tables (Drq .. VfPq, nrJ ..) are re-laid out once at
 * create: they must not change afterwards.  Problems with equal N may share
 * operator arrays. */
typedef struct wadges_group {
  int N, Nq, Nfq;   /* kernel instance, as for wadges_plan_create */
  int K;            /* elements (0 allowed) */
  double dt;        /* this problem's time step */
  const double *vgeo, *vfgeo, *fgeo;
  const uint32_t *mapPq;
  const double *nrJ, *nsJ, *ntJ, *Drq, *Dsq, *Dtq, *VqLq, *VqPq, *VfPq, *Jq;
  double *Q, *rhsQ, *resQ;
  double *Qf[2];
} wadges_group;

typedef struct wadges_batch wadges_batch;

/* Launch shape of the ensemble kernel (the reference's per-kernel
 * elements-per-workgroup knobs KblkV/S/U, euler.c:342-345, as one object that
 * the autotuner chooses).  A tile is the elements one 64-wide wave holds
 * (floor(64/Nq), or floor(64/ceil(Nq/2)) with two rows per lane).  grid = 0:
 * the flat launch, one tile per wave and waves_per_group (1, 2 or 4) waves per
 * workgroup, all of one problem (no loop).  grid > 0: a persistent launch of
 * grid one-wave workgroups over items of tiles_per_item tiles, taken in a
 * static stride (dynamic = 0) or from an atomic work counter (dynamic = 1).
 * order = 1 issues the costliest items first, 0 keeps problem order. */
typedef struct wadges_batch_config {
  int tiles_per_item;
  int grid;
  int dynamic;
  int order;
  int waves_per_group;
} wadges_batch_config;

/* Copies the descriptors, checks them (instances exist, pointers set, Qf[0] !=
 * Qf[1]), allocates the device tables and prepares the operator tables on
 * stream.  The default launch shape is used until wadges_batch_tune or
 * wadges_batch_set_config. */
int wadges_batch_create(wadges_batch **batch, int dim, int ngroups, const wadges_group *groups,
                        double gamma, double tau, void *stream);
int wadges_batch_destroy(wadges_batch *batch);

/* One stage fa, fb for every problem (fdt = its dt); flips the current trace. */
int wadges_batch_stage(wadges_batch *batch, double fa, double fb, void *stream);

/* nsteps LSRK45 steps (five stages each, coefficients of euler.c:2779-2790).
 * use_graph = 1 replays a HIP graph of one step (captured at the first call and
 * after any change of launch shape); 0 launches the stages directly. */
int wadges_batch_run(wadges_batch *batch, int nsteps, int use_graph, void *stream);

int wadges_batch_current(const wadges_batch *batch);   /* 0 or 1: index into Qf[] */
int wadges_batch_get_config(const wadges_batch *batch, wadges_batch_config *cfg);
int wadges_batch_set_config(wadges_batch *batch, const wadges_batch_config *cfg);

/* Autotuner: times every candidate launch shape over reps stages each (HIP
 * events on stream, synchronising), keeps the fastest and returns its mean
 * stage time in *ms (may be NULL).  The state of every problem is saved first
 * and restored afterwards.  Results are cached per ensemble signature (dim and
 * the sorted (Nq, K) list) for the process, and in the text file named by the
 * environment variable WADGES_TUNE_CACHE when it is set. */
int wadges_batch_tune(wadges_batch *batch, int reps, double *ms, void *stream);

/* Introspection. */
int wadges_last_hip_error(void);          /* hipError_t of the last failing HIP call */
const char *wadges_error_string(int code);  /* text for a WADGES_E* code */
int wadges_instance_count(void);            /* number of compiled (dim, Nq, Nfq) instances */
int wadges_instance(int idx, int *dim, int *N, int *Nq, int *Nfq);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* WADGES_H */
