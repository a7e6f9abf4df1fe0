/*
 * tichi_sph.h - C ABI of the MI355X (gfx950) implementation of Tichi_SPH's
 * per-step particle hot path.
 *
 * The library is a set of hand-written HIP kernels behind plain C entry
 * points.  Every pointer inside TsphObject / TsphControl / workspace is a
 * DEVICE pointer owned by the caller (in the Python host layer: torch
 * tensors); the library never allocates or frees device memory and never
 * synchronises the host.  All entry points enqueue work on `stream`
 * (a hipStream_t passed as void*) and return 0 on success or a negative
 * TSPH_E* code; tsph_last_error() then holds a message (thread-local).
 *
 * Arithmetic type: TsphParams.precision selects float (4) or double (8) for
 * every floating-point array; integer arrays are always int32.
 *
 * Layout (per particle object = one reference `Fluid`, sph_obj.py:8-93):
 *   vector fields are 4-wide: [capacity][4] = (x, y, z, w); z = 0 in 2D.
 *   pos.w holds the solver alias X (mass for DFSPH, rest_volume for VFSPH,
 *   sph_obj.py:65-73) so one 16-byte load fetches position and X.
 *   volume_frac is [capacity][phase_num]; drift_vel is [capacity][phase_num][4].
 *   The other w lanes are not part of the reference's vectors: the fused step
 *   keeps records there for later sweeps (vel_adv.w = X / psi between pass A
 *   and the advection, and vel.w the same between the advection and the
 *   tail, which writes 0; normal.w = +-psi, negative where the 0.99
 *   surface-tension gate fails, when gamma != 0 or TSPH_FLAG_ST_RECORDS is
 *   set; drift_vel.w = phi of that phase, or -1 once flagged, in an FBM round).
 *   The fused step keeps particles in cell order (a counting sort each step);
 *   `pid` holds each slot's reference particle index (creation order).
 */
#ifndef TICHI_SPH_H
#define TICHI_SPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSPH_ABI_VERSION 5
#define TSPH_MAX_PHASE 8
#define TSPH_MAX_ITER_LOG 256

enum {
  TSPH_OK = 0,
  TSPH_EINVAL = -1,   /* bad argument / shape */
  TSPH_ECAPACITY = -2,/* capacity or workspace too small */
  TSPH_EHIP = -3,     /* HIP runtime error (message has the HIP string) */
  TSPH_EUNSUPPORTED = -4
};

/* Solver types, sph_config.py:79 / sph_obj.py:65-73. */
enum { TSPH_DFSPH = 0, TSPH_VFSPH = 1 };

/*
 * Per-scene constants.  Replaces the reference's `Config` singleton
 * (sph_config.py:83-232): the host layer fills it from config.json and the
 * scenario's sim_env with the same f64->f32 rounding points.
 */
/* TsphParams.flags */
#define TSPH_FLAG_FAST_QUOTIENT 1 /* FP32 fluid pass A may use the reduced sqrt and r/h
                                     sequences: set only after tsph_verify_fast_quotient
                                     found no difference for these parameters */
#define TSPH_FLAG_ST_RECORDS 2    /* tsph_advect records +-psi (negative where the 0.99
                                     gate fails) in the normal's spare lane also when
                                     gamma = 0, for a surface-tension reverse that
                                     takes d/d gamma at gamma = 0 (with gamma != 0 it
                                     always does); otherwise that lane is 0 */

typedef struct TsphParams {
  int32_t dim, phase_num, solver_type, precision;
  int32_t neighb_range, ntemp, node_num, list_cap_f;
  int32_t grid_size[4], coder[4];
  int32_t cs, wc_gamma, iter_min, iter_max;
  int32_t list_cap_b, flags, pad0, pad1;
  double lb[4], rt[4], gravity[4];
  double phase_rest_density[TSPH_MAX_PHASE];
  double part_size[5], kernel_h[5], kernel_sig[4];
  double r2_max; /* largest r^2 whose fl(fl(sqrt(r^2))/h) < 1 in this precision */
  double dynamic_viscosity, surface_tension_gamma, fbm_diffusion, fbm_convection, cfl_factor;
  double divergence_threshold, compression_threshold;
  /* Per-phase dynamic viscosity, an extension (the reference has one global
   * mu, sph_config.py:148): with mu_per_phase != 0 the viscous force of
   * particle i uses mu_i = sum_k volume_frac_ik * phase_viscosity[k]
   * (k ascending, each term rounded once) in place of dynamic_viscosity. */
  double phase_viscosity[TSPH_MAX_PHASE];
  int32_t mu_per_phase, pad2;
} TsphParams;

/*
 * Device pointers of one particle object (the reference `Fluid`,
 * sph_obj.py:8-93).  Names follow the reference attributes.
 */
typedef struct TsphObject {
  int32_t n, capacity, uid;
  int32_t flags; /* TSPH_OBJ_*: state the library keeps in the descriptor; zero it once */
  void *pos, *vel, *vel_adv, *acce_adv, *acce, *normal, *alpha_1;      /* T[cap][4] */
  void *mass, *rest_density, *rest_volume;                             /* T[cap] */
  void *volume_frac, *volume_frac_tmp;                                 /* T[cap][P] */
  void *drift_vel;                                                     /* T[cap][P][4] */
  void *W, *sph_compression, *sph_density, *psi_adv, *alpha, *alpha_2; /* T[cap] */
  void *kappa, *vol, *fbm_zeta, *pressure;                             /* T[cap] */
  int32_t *pid, *flag, *cell;                 /* [cap]: reference index, FBM flag, cell key */
  int32_t *cell_structured;                   /* [cap][4] neighb_cell_structured_seq */
  int32_t *in_cell;                           /* [cap] neighb_in_cell_seq */
  int32_t *cell_count, *cell_start;           /* [node_num + 2] per-object cell table */
  int32_t *nbr_f, *nbr_b;                     /* [list_cap][cap] slot-major neighbour lists */
  int32_t *nbr_nf, *nbr_nb;                   /* [cap] list lengths (cap+1 = overflow) */
  /* back buffers of the persistent state, swapped by tsph_sort_object */
  void *pos_b, *vel_b, *mass_b, *rest_density_b, *rest_volume_b, *volume_frac_b;
  int32_t *pid_b;
  /* T[cap][4], or null: the boundary's own-object (b,b) sums of W,
   * sph_compression and sph_density in pass A's order.  The boundary is
   * static, so they are kept from one step to the next (TSPH_OBJ_STATIC_SUMS)
   * and each step only adds the (b,f) pairs, with the same result bit for
   * bit.  A boundary with a null pointer is summed in full every step. */
  void *static_sums;
} TsphObject;

/* TsphObject.flags: static_sums holds this object's current (b,b) sums.
 * tsph_density_alpha sets it after summing them; tsph_sort_object clears it,
 * and so must a caller that writes the object's pos, mass or rest_volume. */
#define TSPH_OBJ_STATIC_SUMS 1

/*
 * Device-resident control block: the reference's 0-D fields
 * (config.dt, fluid.compression, statistics_*, general_flag) plus the
 * solver's device-side stop state.  Allocate tsph_control_bytes() bytes of
 * device memory, zero it once, then pass it to every call.
 */
typedef struct TsphControl {
  double dt;             /* config.dt (sph.py:41-47) */
  double compression;    /* fluid.compression (sph.py:219-226); persists across solves */
  int32_t it, stop, counter, general_flag;
  int32_t div_iter, incom_iter, fbm_rounds, solve;
  double ke, pe;         /* statistics_kinetic/gravity_potential_energy (sph.py:272-279) */
  int32_t run_u[TSPH_MAX_ITER_LOG];
  double err_div[TSPH_MAX_ITER_LOG];
  double err_incom[TSPH_MAX_ITER_LOG];
  int32_t overflow_f, overflow_b, counter2, pad;
  uint64_t signal;       /* host address for the early stop signal, or 0 (tsph_solver_begin_signal) */
  int32_t token, pad2;   /* the value the signal's first word receives */
} TsphControl;

const char* tsph_last_error(void);
int tsph_abi_version(void);
size_t tsph_control_bytes(void);

/*
 * Sweeps every FP32 r^2 in [0, r2_max] and writes into mismatches[3] (device
 * int32, overwritten): [0] the inputs for which pass A's reduced sqrt and
 * r / h sequences give a different W(q) or r > 0 than the IEEE operations,
 * for these parameters' h; [1] those with a different W'(q); [2] the largest
 * r^2 bit pattern counted in [0].  No counterpart in the reference: it is the
 * check behind TSPH_FLAG_FAST_QUOTIENT (set it only when [0] is 0), which
 * keeps pass A's strict sums bit-identical.  FP64 parameters write 1, 1, 0
 * (the flag is FP32 only).
 */
int tsph_verify_fast_quotient(const TsphParams* p, int32_t* mismatches, void* stream);

/* Workspace needed by the calls below for an object of `capacity` slots.
 * It carries records between the calls of one step: tsph_density_alpha
 * writes the boundary neighbours' share of alpha_1 that the step's psi
 * half-steps read, and each psi half-step writes the (x, kappa) records its
 * update half-step gathers.  Pass one workspace to every call of a step and
 * run nothing else on it in between (tsph_adj_advect uses both record
 * regions as scratch). */
size_t tsph_workspace_bytes(const TsphParams* p, int32_t capacity);

/* ----------------------------------------------------------------------
 * Neighbour grid
 * ---------------------------------------------------------------------- */

/*
 * Replaces Ngrid.clear_node/encode/mem_shift/fill_node (sph_obj.py:295-323)
 * for the fused step: computes every particle's cell key (node_encode +
 * dim_encode, sph_util.py:65-72), counts particles per cell, prefix-sums the
 * counts and moves the particles into cell order with ascending `pid`
 * inside each cell (the reference's single-thread bucket order).  Particles
 * with an invalid code (0 < code < node_num fails, sph_obj.py:305) go to a
 * trailing bucket.  Swaps the object's front/back state pointers in *obj.
 * When `cfl` is non-zero it also performs cfl_condition (sph.py:41-47) into
 * ctrl->dt from the object's velocities.
 */
int tsph_sort_object(const TsphParams* p, TsphObject* obj, TsphControl* ctrl, int32_t cfl,
                     void* workspace, size_t ws_bytes, void* stream);

/*
 * tsph_sort_object without a workspace.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory or pass a workspace
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * allocation).  The sort's scratch (keys, arrival ranks, provisional slots,
 * scan sums, CFL partials) lives in a file-scope device array of
 * tsph_sort_static_scratch_bytes() bytes (about 42 MB) that each device
 * holds from the library's first use on it until the process ends; its
 * address is looked up once per device, on the first call there.  The call
 * itself allocates nothing.  Limits: obj->n <= tsph_sort_static_max_particles()
 * (2^21) and node_num <= tsph_sort_static_max_cells() (2^24); larger inputs
 * return TSPH_ECAPACITY without launching anything.  Every call on a device
 * shares that device's array, so calls that overlap on one device (other
 * streams, other host threads) race; calls ordered on one stream do not.
 * Results are bit-identical to tsph_sort_object's.
 */
int tsph_sort_object_static(const TsphParams* p, TsphObject* obj, TsphControl* ctrl, int32_t cfl,
                            void* stream);
size_t tsph_sort_static_scratch_bytes(void);
int32_t tsph_sort_static_max_particles(void);
int32_t tsph_sort_static_max_cells(void);

/*
 * The reference's shared Ngrid arrays (sph_obj.py:283-293), built over the
 * objects' current storage order: node_part_count/shift/shift_count per cell
 * and part_pid_in_node/part_uid_in_node per slot, fluid entries before
 * boundary entries in a cell and ascending reference index within each
 * (the single-thread order of main.py:178-183).  Also writes each object's
 * neighb_cell_seq (cell), neighb_cell_structured_seq and neighb_in_cell_seq.
 * part_pid_in_node holds storage slots; objs[u] must have uid u (1 or 2
 * objects).  Used by the reference-surface
 * kernels (tsph_ref_kernel) and for parity export.
 */
int tsph_ngrid_build(const TsphParams* p, TsphObject* const* objs, int32_t nobj,
                     int32_t* node_part_count, int32_t* node_part_shift,
                     int32_t* node_part_shift_count, int32_t* part_pid_in_node,
                     int32_t* part_uid_in_node, void* workspace, size_t ws_bytes, void* stream);

/*
 * NeighborSearch.fill_neighbs (NeighborSearch.py:84-101) over a built Ngrid:
 * for each particle of `obj`, the unfiltered candidate list (self included)
 * grouped by uid, template order then cell order.  Pass neighbs == NULL to
 * get only neighbs_count (per particle, per uid) from which the caller sizes
 * neighbs and computes row starts; then call again with neighbs and
 * row_start (int32[n+1]).  neighbs_shift/count are [n][n_uids].
 */
int tsph_neighbor_lists(const TsphParams* p, const TsphObject* obj, int32_t n_uids,
                        const int32_t* node_part_count, const int32_t* node_part_shift,
                        const int32_t* part_pid_in_node, const int32_t* part_uid_in_node,
                        const int32_t* row_start, int32_t* neighbs, int32_t* neighbs_shift,
                        int32_t* neighbs_count, void* stream);

/* ----------------------------------------------------------------------
 * Fused per-step kernels (sph_step, main.py:175-256).  Objects must have
 * been sorted by tsph_sort_object in this step.
 * ---------------------------------------------------------------------- */

/*
 * SPH_clean_value + SPH_prepare_attr (f,f)(f,b)(b,b)(b,f) + SPH_prepare_alpha_1
 * (f,f)(f,b) + SPH_prepare_alpha_2 (f,f)(b,f) + SPH_prepare_alpha (f)(b) +
 * SPH_vel_2_vel_adv (sph.py:22-105, 253-256; main.py:185-200), in one sweep
 * per particle.  Records each fluid particle's in-range fluid and boundary
 * neighbours (0 < r < h) in reference order for the later sweeps, and opens
 * the divergence solve's stop state.
 */
int tsph_density_alpha(const TsphParams* p, TsphObject* fluid, TsphObject* bound,
                       TsphControl* ctrl, void* workspace, size_t ws_bytes, void* stream);

/*
 * Opens a Jacobi solve (main.py:199-201 / 221-222): it = 0 and the entry test
 * of the while loop.  solve = 0 divergence, 1 density.
 */
int tsph_solver_begin(const TsphParams* p, TsphControl* ctrl, int32_t solve, void* stream);

/*
 * tsph_solver_begin plus an early stop signal for the host.  This has no
 * counterpart in the reference, whose host loop reads the stop flag back
 * after every iteration (main.py:201-213, 222-234).  `signal` is pinned host
 * memory mapped for the device (one 8-byte aligned uint64).  Every update
 * half-step of this solve that runs stores, in one system-scope 64-bit store,
 * (token << 32) | (2 * it + stop): the iterations done and whether the loop
 * ends there.  The decision is taken when the update launch starts, so a host
 * that waits for its token can queue the work that follows the solve while
 * the last update still runs, instead of draining the stream.  A solve that
 * runs no iteration writes nothing.  signal = NULL behaves as
 * tsph_solver_begin.
 */
int tsph_solver_begin_signal(const TsphParams* p, TsphControl* ctrl, int32_t solve,
                             uint64_t* signal, int32_t token, void* stream);

/* Pinned, device-mapped, coherent host memory for the signal above (`bytes`
 * bytes, zeroed): *host is the address the host reads, *device the one to
 * pass as `signal`.  Release it with tsph_signal_free(host). */
int tsph_signal_alloc(size_t bytes, void** host, void** device);
int tsph_signal_free(void* host);

/*
 * One predicated Jacobi iteration k of a solve: IPPE_adv_psi_init +
 * IPPE_adv_psi (f,f)(f,b) + IPPE_psi_adv_non_negative (sph.py:196-226) with a
 * deterministic device reduction of the compression, then
 * IPPE_update_vel_adv (f,f)(f,b) (sph.py:229-244).  Does nothing once the
 * device stop state says the loop has ended (main.py:201-213, 222-234).
 */
int tsph_solver_iteration(const TsphParams* p, TsphObject* fluid, TsphObject* bound,
                          TsphControl* ctrl, int32_t k, void* workspace, size_t ws_bytes,
                          void* stream);

/* The two halves of tsph_solver_iteration as separate launches:
 * psi = IPPE_adv_psi_init + IPPE_adv_psi (f,f)(f,b) + IPPE_psi_adv_non_negative
 * (sph.py:196-226) with the stop decision; update = IPPE_update_vel_adv
 * (f,f)(f,b) (sph.py:229-244). */
int tsph_solver_psi(const TsphParams* p, TsphObject* fluid, TsphObject* bound, TsphControl* ctrl,
                    int32_t k, void* workspace, size_t ws_bytes, void* stream);
int tsph_solver_update(const TsphParams* p, TsphObject* fluid, TsphObject* bound,
                       TsphControl* ctrl, int32_t k, void* workspace, size_t ws_bytes,
                       void* stream);

/*
 * SPH_vel_adv_2_vel + SPH_advection_gravity_acc + SPH_advection_viscosity_acc
 * (f,f) + SPH_advection_surface_tension_acc (f,f) + SPH_advection_update_vel_adv
 * (sph.py:108-166, 247-262; main.py:214-219).
 */
int tsph_advect(const TsphParams* p, TsphObject* fluid, TsphObject* bound, TsphControl* ctrl,
                void* workspace, size_t ws_bytes, void* stream);

/*
 * One FBM round (main.py:244-248): SPH_FBM_clean_tmp + SPH_FBM_convect (f,f) +
 * SPH_FBM_diffuse (f,f) + SPH_FBM_check_tmp (sph.py:301-372).  Predicated on
 * ctrl->general_flag; ctrl->fbm_rounds counts rounds that ran.
 */
int tsph_fbm_round(const TsphParams* p, TsphObject* fluid, TsphControl* ctrl, void* workspace,
                   size_t ws_bytes, void* stream);

/*
 * Step tail: [SPH_update_volume_frac] + SPH_update_mass + SPH_update_pos +
 * SPH_update_energy (sph.py:265-286, 375-379; main.py:250-253).  fbm != 0
 * applies the volume-fraction update first.
 */
int tsph_tail(const TsphParams* p, TsphObject* fluid, TsphControl* ctrl, int32_t fbm,
              void* workspace, size_t ws_bytes, void* stream);

/*
 * The WCSPH block of sph_step (main.py:236-240): WC_pressure_val (fluid),
 * WC_pressure_acce (f,f)(f,b) and SPH_advection_update_vel_adv
 * (sph.py:169-193, 247-250), after the density solve, over this step's
 * neighbour lists.
 */
int tsph_wcsph(const TsphParams* p, TsphObject* fluid, TsphObject* bound, TsphControl* ctrl,
               void* stream);

/*
 * map_velocity (sph.py:382-397) onto a Grid (sph_obj.py:328-344): node I of
 * the row-major grid of grid_shape[dim] nodes (last axis fastest) sits at
 * sim_space_lb + I * grid_dist; grid_vel (T[nodes][4]) receives the sum of
 * X_j / sph_psi_j * vel_j * W(|x_I - x_j|) over obj's particles in the cells
 * around the node's cell, in the reference's order.  obj's cell table must
 * be this step's (tsph_sort_object), as the reference's Ngrid is.
 */
int tsph_map_velocity(const TsphParams* p, const TsphObject* obj, const int32_t* grid_shape,
                      double grid_dist, void* grid_vel, void* stream);

/* ----------------------------------------------------------------------
 * Reverse mode (adjoint) of the DFSPH step.  The reference has no
 * autodiff; these entries differentiate the fused step above (sph_step,
 * main.py:175-256) for one phase, DFSPH, fixed dt, no FBM, with the Jacobi
 * iteration counts and the clamp / neighbour sets of the recorded forward
 * step held fixed.  Every reverse sweep is the transpose of a forward sweep
 * written as a gather over the same symmetric neighbour lists, so no
 * atomics are used and results are deterministic.
 * ---------------------------------------------------------------------- */

/* Adjoint buffers of one step, all [capacity] in the step's cell order.
 * V4 arrays hold 4 values of the precision's float type, T arrays one. */
typedef struct TsphAdjoint {
  void *xbar;         /* V4: d loss / d pos at the step's start (accumulated) */
  void *vbar;         /* V4: d loss / d (velocity state being reversed), read */
  void *vbar_out;     /* V4: written by the sweeps that move the velocity state back */
  void *rhobar;       /* T: d loss / d sph_density (accumulated) */
  void *alphabar;     /* T: d loss / d alpha (accumulated) */
  void *nbar;         /* V4: d loss / d normal (NULL: see tsph_adj_advect) */
  void *a1bar;        /* V4: d loss / d alpha_1 (xyz) and d alpha_2 (w) */
  void *tape_vel_adv; /* V4: recorded vel_adv entering the iteration being reversed; its w
                         lane is overwritten (the reverse consumes the record) */
  void *kappa_next;   /* T: recorded kappa (psi_adv / alpha) of the iteration that the next
                         tsph_adj_solver_iteration call reverses, or NULL if none: every
                         call that writes the velocity adjoint stores it in vbar's w lane */
  void *tape_vel;     /* V4: recorded velocity after the advection (its w lane: X / psi) */
  void *tape_normal;  /* V4: recorded surface-tension normals */
  double *grad;       /* [2]: accumulated d loss / d dynamic_viscosity, d surface_tension_gamma */
  void *bmom;         /* V4 [3][capacity]: each fluid particle's boundary sums of the step
                         being reversed (tsph_adj_bound_sums) */
  int32_t *bflag;     /* [capacity]: 1 where a particle has boundary neighbours, else 0 */
} TsphAdjoint;

/* tsph_solver_psi that also records vel_adv entering iteration k and the
 * iteration's kappa into the two [capacity] tape arrays (the forward half of
 * tsph_adj_solver_iteration). */
int tsph_solver_psi_taped(const TsphParams* p, TsphObject* fluid, TsphObject* bound,
                          TsphControl* ctrl, int32_t k, void* tape_vel_adv, void* tape_kappa,
                          void* workspace, size_t ws_bytes, void* stream);

/* tsph_solver_psi with the clamp set of a recorded iteration: psi_adv is
 * set to 0 exactly where kappa_ref (a tape_kappa of tsph_solver_psi_taped,
 * in the same cell order) is 0, and left unclamped elsewhere, instead of
 * clamping the negative values (IPPE_psi_adv_non_negative, sph.py:219-226).
 * It replays the path whose derivative the adjoint takes, so that finite
 * differences of replayed steps do not cross the clamp's kink. */
int tsph_solver_psi_frozen(const TsphParams* p, TsphObject* fluid, TsphObject* bound,
                           TsphControl* ctrl, int32_t k, const void* kappa_ref, void* workspace,
                           size_t ws_bytes, void* stream);

/* The fluid's neighbour lists and counts of tsph_density_alpha alone,
 * rebuilt (bit-identically) from recorded sorted positions and cell table,
 * so that per-step lists are recomputed instead of stored.  Densities and
 * alpha factors are not recomputed: the adjoint keeps them in its tapes. */
int tsph_relist(const TsphParams* p, TsphObject* fluid, TsphObject* bound, TsphControl* ctrl,
                void* workspace, size_t ws_bytes, void* stream);

/* The boundary neighbours' sums that the reverse of the Jacobi iterations
 * and of pass A use in place of the (f,b) pairs (boundary velocities and
 * kappa are zero, so those pairs depend on the boundary only through them):
 * for each fluid particle i, over the boundary particles b in range
 * (0 < r^2 <= r2_max, the lists' test), with x = x_i - x_b, X_b = pos.w,
 * g = W'(r)/r and hq = (W'' - g)/r^2:
 *   bmom[i]            = (sum X_b g x, sum X_b g)
 *   bmom[capacity + i] = sum X_b hq (xx, yy, zz, xy)
 *   bmom[2capacity+i]  = sum X_b hq (xz, yz, 0, 0)
 * and bflag[i] = 1 if any b is in range.  Walks the boundary cells (not the
 * lists) of the particles whose boundary-list count nbr_nb is not 0, so the
 * fluid descriptor must hold the step's counts (as tsph_density_alpha or
 * tsph_relist left them).  Call it once per reversed step, before
 * tsph_adj_begin. */
int tsph_adj_bound_sums(const TsphParams* p, TsphObject* fluid, TsphObject* bound,
                        TsphAdjoint* adj, void* stream);

/* Opens the reverse of one step (tail, SPH_update_pos sph.py:265-269):
 * reads d loss / d pos and d loss / d vel after the step from V4 arrays in
 * reference (pid) order, gathers them into the step's cell order with
 * pid[], and clears the accumulated density / alpha adjoints. */
int tsph_adj_begin(const TsphParams* p, TsphObject* fluid, TsphControl* ctrl, TsphAdjoint* adj,
                   const int32_t* pid, const void* xbar_ref, const void* vbar_ref, void* stream);

/* Reverse of one Jacobi iteration (IPPE_adv_psi + IPPE_psi_adv_non_negative
 * + IPPE_update_vel_adv, sph.py:196-244) whose record is adj->tape_vel_adv,
 * with that iteration's kappa in vbar's w lane: vbar -> vbar_out,
 * accumulating xbar, rhobar, alphabar.  The (f,b) pairs come from the step's
 * boundary sums (tsph_adj_bound_sums); `bound` is not read. */
int tsph_adj_solver_iteration(const TsphParams* p, TsphObject* fluid, TsphObject* bound,
                              TsphControl* ctrl, TsphAdjoint* adj, void* workspace,
                              size_t ws_bytes, void* stream);

/* Reverse of the advection (viscosity, normals, surface tension and
 * vel_adv = vel + dt * acce_adv; sph.py:108-166, 247-250): vbar -> vbar_out,
 * accumulating xbar, rhobar and adj->grad.  adj->nbar may be NULL when
 * surface_tension_gamma == 0 and d loss / d gamma is not wanted: every
 * surface-tension term then vanishes and that reverse sweep is skipped.
 * The workspace's record regions hold per-particle records of the reverse
 * sweeps during the call (see tsph_workspace_bytes). */
int tsph_adj_advect(const TsphParams* p, TsphObject* fluid, TsphControl* ctrl, TsphAdjoint* adj,
                    void* workspace, size_t ws_bytes, void* stream);

/* Reverse of density and alpha (SPH_prepare_attr, _alpha_1, _alpha_2,
 * _alpha; sph.py:50-105): accumulates xbar from rhobar and alphabar.  The
 * (f,b) pairs come from the step's boundary sums (tsph_adj_bound_sums). */
int tsph_adj_density_alpha(const TsphParams* p, TsphObject* fluid, TsphObject* bound,
                           TsphAdjoint* adj, void* stream);

/* Closes the reverse of one step: scatters xbar and vbar (d loss / d pos and
 * d loss / d vel at the step's start) to reference order with pid[]. */
int tsph_adj_end(const TsphParams* p, TsphObject* fluid, TsphAdjoint* adj, const int32_t* pid,
                 void* xbar_ref, void* vbar_ref, void* stream);

/* ----------------------------------------------------------------------
 * Batched scenes (no reference counterpart: the reference runs one scene
 * per process).  Many independent scenes of uneven particle counts are
 * advanced together: every call below is ONE grouped launch per kernel over
 * all scenes, driven by a device-resident scene table (constants, object
 * descriptors, control block and workspace of each scene).  Each scene keeps
 * its own stop rule; scenes whose solve has ended are predicated off on the
 * device.  The table changes only on the device (the sort swaps each
 * scene's front/back state pointers inside it), so a captured hipGraph of
 * these calls stays valid across replays.  Particle counts must not change
 * while a table is in use.  FBM is not batched.
 * ---------------------------------------------------------------------- */
typedef struct TsphBatchInfo {
  int32_t nscenes, precision;
  int32_t tiles[4]; /* tiles of a grouped launch: fluid, boundary, cell, scene */
} TsphBatchInfo;

/* Device bytes of the scene table for nscenes scenes. */
size_t tsph_batch_table_bytes(int32_t precision, int32_t nscenes);

/* Builds the scene table for scenes s = 0..nscenes-1 (params[s], fluid[s],
 * bound[s], ctrl[s], workspace[s] of ws_bytes[s]) in `host_staging` (pinned
 * host memory of table_bytes, which must stay untouched until the stream
 * has executed the copy) and copies it to the device `table`.  Boundary
 * objects must already be sorted.  Fills *info. */
int tsph_batch_prepare(const TsphParams* params, const TsphObject* fluid, const TsphObject* bound,
                       TsphControl* const* ctrl, void* const* workspace, const size_t* ws_bytes,
                       int32_t nscenes, void* host_staging, void* table, size_t table_bytes,
                       TsphBatchInfo* info, void* stream);

/* tsph_sort_object for every scene's fluid.  Swaps the front/back pointers
 * inside the device table and in the caller's host copies fluid[0..n-1]
 * (pass NULL to leave host copies alone, e.g. when replaying a graph). */
int tsph_batch_sort(const TsphBatchInfo* info, void* table, TsphObject* fluid, int32_t cfl,
                    int32_t grid, void* stream);
/* tsph_density_alpha, tsph_solver_begin, _psi, _update, tsph_advect and
 * tsph_tail (fbm = 0) for every scene.  `grid` = number of blocks of the
 * grouped launch (0: one block per tile); the result does not depend on it.
 * tsph_batch_density_alpha sets TSPH_OBJ_STATIC_SUMS in the table's boundary
 * descriptors, as tsph_density_alpha does in the caller's: a binding that
 * then writes a boundary's positions or masses prepares the table again from
 * descriptors with the flag cleared. */
int tsph_batch_density_alpha(const TsphBatchInfo* info, void* table, int32_t grid, void* stream);
int tsph_batch_solver_begin(const TsphBatchInfo* info, void* table, int32_t solve, void* stream);
/* tsph_batch_solver_begin with each scene's early stop signal (as
 * tsph_solver_begin_signal, no reference counterpart): scene s's update
 * half-steps store (token << 32) | (2 * it + stop) into signal[s], an array
 * of nscenes uint64 in pinned, device-mapped host memory (tsph_signal_alloc),
 * so that the host learns when every scene's solve has ended without
 * draining the stream. */
int tsph_batch_solver_begin_signal(const TsphBatchInfo* info, void* table, int32_t solve,
                                   uint64_t* signal, int32_t token, void* stream);
int tsph_batch_solver_psi(const TsphBatchInfo* info, void* table, int32_t k, int32_t grid,
                          void* stream);
int tsph_batch_solver_update(const TsphBatchInfo* info, void* table, int32_t k, int32_t grid,
                             void* stream);
int tsph_batch_advect(const TsphBatchInfo* info, void* table, int32_t grid, void* stream);
int tsph_batch_tail(const TsphBatchInfo* info, void* table, int32_t grid, void* stream);
/* Gathers every scene's (it, stop) of its current solve into the device
 * int32 array out[2 * nscenes], for one copy to the host. */
int tsph_batch_stop_state(const TsphBatchInfo* info, void* table, int32_t* out, void* stream);

/* ----------------------------------------------------------------------
 * Reference kernel surface: the sph.py kernels one by one, with the
 * reference's (ngrid, obj, nobj) signature, walking the Ngrid built by
 * tsph_ngrid_build exactly as the reference does.  `kernel` is one of the
 * TSPH_K_* ids; nobj may be NULL for per-particle kernels.
 * ---------------------------------------------------------------------- */
enum {
  TSPH_K_SPH_clean_value = 0,                 /* sph.py:22-38 */
  TSPH_K_cfl_condition = 1,                   /* sph.py:41-47 */
  TSPH_K_SPH_prepare_attr = 2,                /* sph.py:50-64 */
  TSPH_K_SPH_prepare_alpha_1 = 3,             /* sph.py:67-81 */
  TSPH_K_SPH_prepare_alpha_2 = 4,             /* sph.py:84-97 */
  TSPH_K_SPH_prepare_alpha = 5,               /* sph.py:100-105 */
  TSPH_K_SPH_advection_gravity_acc = 6,       /* sph.py:108-111 */
  TSPH_K_SPH_advection_viscosity_acc = 7,     /* sph.py:114-129 */
  TSPH_K_SPH_advection_surface_tension_acc = 8, /* sph.py:132-166 */
  TSPH_K_WC_pressure_val = 9,                 /* sph.py:169-175 */
  TSPH_K_WC_pressure_acce = 10,               /* sph.py:178-193 */
  TSPH_K_IPPE_adv_psi_init = 11,              /* sph.py:196-199 */
  TSPH_K_IPPE_adv_psi = 12,                   /* sph.py:202-216 */
  TSPH_K_IPPE_psi_adv_non_negative = 13,      /* sph.py:219-226 */
  TSPH_K_IPPE_update_vel_adv = 14,            /* sph.py:229-244 */
  TSPH_K_SPH_advection_update_vel_adv = 15,   /* sph.py:247-250 */
  TSPH_K_SPH_vel_2_vel_adv = 16,              /* sph.py:253-256 */
  TSPH_K_SPH_vel_adv_2_vel = 17,              /* sph.py:259-262 */
  TSPH_K_SPH_update_pos = 18,                 /* sph.py:265-269 */
  TSPH_K_SPH_update_energy = 19,              /* sph.py:272-279 */
  TSPH_K_SPH_update_mass = 20,                /* sph.py:282-286 */
  TSPH_K_SPH_FBM_clean_tmp = 21,              /* sph.py:301-305 */
  TSPH_K_SPH_FBM_diffuse = 22,                /* sph.py:308-326 */
  TSPH_K_SPH_FBM_convect = 23,                /* sph.py:329-363 */
  TSPH_K_SPH_FBM_check_tmp = 24,              /* sph.py:366-372 */
  TSPH_K_SPH_update_volume_frac = 25,         /* sph.py:375-379 */
  TSPH_K_COUNT = 26
};

int tsph_ref_kernel(const TsphParams* p, int32_t kernel, TsphObject* obj, const TsphObject* nobj,
                    const int32_t* node_part_count, const int32_t* node_part_shift,
                    const int32_t* part_pid_in_node, const int32_t* part_uid_in_node,
                    TsphControl* ctrl, void* workspace, size_t ws_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TICHI_SPH_H */
