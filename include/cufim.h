/*
 * cufim.h -- C ABI of the MI355X FIM activation-time solver (libcufim.so).
 *
 * Drop-in boundary for the reference activation-time path.  The reference
 * exposes that path as a C++ class, not an FFI:
 *   tw::CudaSolverFIM(const ek_data&)            src/eikonal/cuda_fim.h:1465
 *   double CudaSolverFIM::solve(act, apd, file, reps)
 *                                                src/eikonal/cuda_fim.h:1468-1470,
 *                                                src/eikonal/cuda_fim.cu:2386-2503
 * The entry points below are what that class does, split at the host/device
 * line so any host language can bind them with plain pointers and sizes.
 * The C++ mirror of the class (fim::SolverFIM) and the reference-format CLI
 * are built on top of this header.
 *
 * Conventions
 *   - every int-returning call returns 0 on success and a negative FIM_E*
 *     code on failure; fim_last_error() gives a message (thread-local).
 *     Nothing here calls exit() or abort(), and no C++ exception crosses it.
 *   - units are the reference's: micrometres, micrometres per microsecond
 *     (= m/s), microseconds.
 *   - calls on one fim_ctx must be serialised by the caller; distinct
 *     contexts are independent.  Host objects (fim_mesh, fim_pack) are
 *     read-only after construction and may be shared across threads.
 */
#ifndef CUFIM_H_
#define CUFIM_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FIM_OK 0
#define FIM_EINVAL (-1)   /* bad argument or inconsistent input          */
#define FIM_EIO (-2)      /* file not found / unreadable / bad format    */
#define FIM_ELIMIT (-3)   /* a size limit was exceeded (LDS, u16 ids)    */
#define FIM_EDEVICE (-4)  /* HIP runtime error                           */
#define FIM_ENOCONV (-5)  /* iteration cap reached before convergence    */
#define FIM_ESTATE (-6)   /* call not valid in the context's mode/state  */

/* vertex states, numbering of the reference enum ek_state
 * (src/core/types.h:85) */
#define FIM_FAR 0
#define FIM_NEIG 1
#define FIM_NARROW 2
#define FIM_NEXT 3
#define FIM_FROZEN 4
#define FIM_TAIL 5

/* solver modes */
#define FIM_MODE_CONVERGE 0 /* plain FIM to the fixed point, FP32, exact local
                               solver (tet + tri + edge); batched instances  */
#define FIM_MODE_COMPAT 1   /* the reference time-window solver, FP64, with the
                               reference local update as written           */

/* local solver variants */
#define FIM_LOCAL_EXACT 0 /* minimum over the opposite face (tet, tri, edge) */
#define FIM_LOCAL_REF 1   /* reference update_per_element
                             (src/eikonal/cuda_fim.h:1362-1411)            */

/* partition sources for fim_pack_build */
#define FIM_PART_TAGS 0 /* element tags as given (reference .tags input,
                           reference submesh numbering, types.cpp:306-570)  */
#define FIM_PART_RCB 1  /* own seeded recursive bisection, LDS-budgeted,
                           submesh-contiguous renumbering                  */

const char *fim_last_error(void);

/* ========================================================================
 * Host-side mesh (replaces tw::mesh_t + load_vtk + read_vector_txt +
 * compute_base_connectivity, src/core/types.h:221-254, src/io/vtk.cpp:11-105,
 * src/io/io_base.h:82-86, src/core/types.cpp:43-225)
 * ===================================================================== */
typedef struct fim_mesh fim_mesh;

/* Build from arrays (copied).  e2n: 4*ne vertex ids (any order; sorted per
 * element as the reference does).  fib: 3*ne unit fibre per element or NULL
 * for (1,0,0).  vel: 2*ne (v_fibre, v_cross) per element or NULL for 0.6.
 * etags: ne element tags or NULL (all 1). */
int fim_mesh_create(int64_t nn, int64_t ne, const double *xyz, const int64_t *e2n,
                    const double *fib, const double *vel, const int32_t *etags,
                    fim_mesh **out);
/* legacy ASCII VTK tetrahedral mesh (reference load_vtk semantics: fibres
 * (1,0,0), speed 0.6) */
int fim_mesh_load_vtk(const char *path, fim_mesh **out);
/* partition tags, one integer per element (reference .tags file) */
int fim_mesh_read_tags(fim_mesh *m, const char *path);
int fim_mesh_set_tags(fim_mesh *m, const int32_t *atags);
/* per-element region ids 0..3 (regional conduction-velocity scalings) */
int fim_mesh_set_regions(fim_mesh *m, const uint8_t *region);
int fim_mesh_info(const fim_mesh *m, int64_t *nn, int64_t *ne);
/* copies out xyz (3*nn), element nodes sorted per element (4*ne), element
 * tags (ne); any pointer may be NULL */
int fim_mesh_get(const fim_mesh *m, double *xyz, int64_t *e2n_sorted, int32_t *etags);
/* per-element fibre (3*ne) and speeds (2*ne: v_fibre, v_cross); either may
 * be NULL */
int fim_mesh_get_material(const fim_mesh *m, double *fib, double *vel);
/* reference front width 1.01 * max edge length / 0.6
 * (src/eikonal/cuda_fim.h:1536-1576) */
double fim_mesh_front_width(const fim_mesh *m);
/* write the reference VTK output (src/io/vtk.cpp:109-212): points %.3f,
 * CELL_TYPES before CELLS, sorted element rows, elemTag, and POINT_DATA
 * phi / states (values >= 1e50 written as -1).  phi/state may be NULL. */
int fim_mesh_save_vtk(const fim_mesh *m, const char *path, const double *phi,
                      const int32_t *state);
/* Conduction velocities from a plan: every element whose tag is listed by a
 * region (plan "config" -> "functions"[f].CV.CVref) gets v_fibre = vf and
 * v_cross = vs (the metric is transversely isotropic; vn is not used).  The
 * reference parses these (src/io/io_json.cpp:82-128) but solves with a fixed
 * 0.6 m/s (src/core/types.cpp:231); this entry point is opt-in and the
 * drop-in path does not call it.  n_set = number of elements changed. */
int fim_mesh_apply_plan_speeds(fim_mesh *m, const char *plan_path, int64_t *n_set);
void fim_mesh_destroy(fim_mesh *m);

/* Synthetic meshes (own generators, seeded, deterministic).
 * Box [0,L]^3 split into nx*ny*nz hexes x 6 tets (Kuhn split along the main
 * diagonal), interior vertices jittered uniformly by +-jitter*spacing per
 * axis (SplitMix64-seeded xoshiro256**).  Fibre angle rotates linearly in z
 * from angle0 to angle1 (degrees, in the xy plane); speeds vf, vc. */
int fim_mesh_gen_box(double lx, double ly, double lz, int nx, int ny, int nz,
                     double jitter, uint64_t seed, double angle0, double angle1,
                     double vf, double vc, fim_mesh **out);

/* Prolate-ellipsoid shell: hexes of side h (um) whose centre lies between
 * the inner (semi-axes ai, bi, bi) and outer (ao, bo, bo) ellipsoids about the
 * origin, 6 tets each (same split as the box).  Fibres rotate transmurally
 * from +60 deg (inner) to -60 deg (outer) in the tangent plane, from the
 * circumferential direction about the x axis towards the longitudinal one;
 * speeds vf / vc, and vs (both) inside n_scars seeded spheres of radius
 * r0..r1 centred at random kept hexes.  Region id 1 marks scar elements. */
int fim_mesh_gen_shell(double ao, double bo, double ai, double bi, double h, int32_t n_scars,
                       double r0, double r1, uint64_t seed, double vf, double vc, double vs,
                       fim_mesh **out);

/* Disjoint union of meshes (vertices and elements renumbered by offsets, in
 * argument order; tags, fibres, speeds and regions carried over).  One
 * context on the union batches solves of different meshes: an instance
 * seeded in one component never activates the others. */
int fim_mesh_concat(const fim_mesh *const *parts, int32_t n, fim_mesh **out);

/* ========================================================================
 * Pack: the submesh decomposition in device layout (replaces
 * extract_submeshes2 + smsh_t/eksv_t/inter3_t blobs, src/core/types.cpp:
 * 306-570, src/eikonal/cuda_fim.h:752-908, 1196-1239, cuda_fim.cu:1218-1501)
 * ===================================================================== */
typedef struct fim_pack fim_pack;

typedef struct {
  int32_t source;   /* FIM_PART_TAGS or FIM_PART_RCB                      */
  int32_t max_nn;   /* RCB: max vertex copies per submesh (e.g. 512)       */
  int32_t max_ne;   /* RCB: max elements per submesh (0 = derived)         */
  uint64_t seed;    /* RCB: tie-break seed (the partition is a pure
                       function of mesh and seed)                          */
} fim_part_opts;

typedef struct {
  int64_t nn_global, ne_global;
  int64_t nsub, ncopy, n_iface_copies;
  int32_t nn_max, ne_max;
  int64_t device_bytes_mesh;  /* compact FP32 mesh bytes (plain mode)      */
  int64_t lds_bytes;          /* dynamic LDS per workgroup (plain mode)    */
} fim_pack_info_t;

int fim_pack_build(const fim_mesh *m, const fim_part_opts *o, fim_pack **out);
int fim_pack_info(const fim_pack *p, fim_pack_info_t *info);
/* element -> submesh id (ne) and local->global vertex ids per copy (ncopy) */
int fim_pack_partition(const fim_pack *p, int64_t *e2s, int64_t *c2g);
void fim_pack_destroy(fim_pack *p);
/* writes the element -> submesh map as a tags file (one integer per line,
 * the format fim_mesh_read_tags and the reference's --part read), so a
 * partition made here can be fed to both (no counterpart in the reference,
 * which needs an external METIS run, src/main.cpp:182) */
int fim_pack_save_tags(const fim_pack *p, const char *path);

/* ========================================================================
 * Device solver context (replaces CudaSolverFIM's device side)
 * ===================================================================== */
typedef struct fim_ctx fim_ctx;

typedef struct {
  int32_t mode;          /* FIM_MODE_*                                       */
  int32_t local;         /* FIM_LOCAL_* (converge mode; compat uses REF
                            unless set to EXACT)                            */
  int32_t n_instances;   /* independent solves on the pack (converge mode)   */
  int32_t device;        /* HIP device ordinal                               */
  int32_t block;         /* workgroup size: 64, 128, 256, 512 (0 = 512 in    */
                         /* converge mode, 256 in compat mode)               */
  int32_t local_iters;   /* max local iterations per activation (0 = until
                            locally converged)                              */
  int32_t graph_iters;   /* global iterations per captured graph, rounded up
                            to a multiple of 3 (0 = 6)                      */
  int32_t grid;          /* workgroups of the persistent kernels (0 = auto)  */
  double eps_rel;        /* convergence threshold (0 = 1e-6, reference)      */
  double apd;            /* compat: action potential duration, us (0 = 3e5)  */
  int32_t profile;       /* converge: launch iterations without the graph and
                            time every compute-kernel launch with events    */
  int32_t staged;        /* converge: 1 = copy each activated submesh's
                            elements into LDS and scan them every sweep;
                            0 = elements stay in global memory and the
                            active ones are found from flagged vertices     */
  double window;         /* converge: time window (us) added per global
                            iteration; a flagged vertex later than
                            (iteration+1)*window is lowered but does not act
                            as a source until the window passes it (0 = off,
                            plain FIM).  The fixed point does not depend on
                            it.                                             */
  int32_t unit;          /* converge: queued instances of one submesh that
                            one workgroup runs back to back as a work unit
                            (0 = 2); the units of a list are dealt out per
                            XCD so a submesh's element data stays in one L2 */
  int32_t waves;         /* converge, workgroup 256: waves per SIMD the
                            compute kernel's registers are sized for, 5..8
                            (0 = as many as the LDS carve admits).  Fewer,
                            spill-free waves suit lists too short to fill
                            the chip (single solves).                       */
} fim_opts;

typedef struct {
  int64_t global_iters;     /* global iterations run by the last fim_run    */
  int64_t activations;      /* submesh activations (work items)             */
  int64_t local_iters;      /* local iterations summed over activations     */
  int64_t elem_updates;     /* element evaluations                          */
  double last_ms;           /* device time of the last fim_run / fim_step   */
  int64_t alg_bytes;        /* algorithmic HBM bytes of the activations:
                               32 per distinct element evaluated (FP32
                               geometry + u16 node ids) + 10 per vertex
                               copy + 12 per interface copy of each
                               activated submesh                            */
  int64_t compute_launches; /* compute-kernel launches timed (profile)      */
  double compute_ms;        /* summed compute-kernel time (profile)         */
  int64_t vertex_updates;   /* (element, target vertex) candidates computed */
} fim_stats_t;

int fim_create(const fim_pack *p, const fim_opts *o, fim_ctx **out);
/* reset_states_gpu (src/eikonal/cuda_fim.cu:479-542): every copy of every
 * instance to far / +inf */
int fim_reset(fim_ctx *c);
/* seed stimulus vertices (global ids) of one instance with time t0 on all
 * their copies; update_events_gpu (src/eikonal/cuda_fim.cu:632-683) */
int fim_seed(fim_ctx *c, int32_t inst, const int64_t *vtx, int64_t n, double t0);
/* fim_seed for many instances in one transfer and one launch: instance
 * inst[i] gets vertices vtx[off[i] .. off[i+1]) (off has n_inst+1 entries,
 * off[0] = 0).  No counterpart in the reference (one solve per process);
 * for batched solves. */
int fim_seed_batch(fim_ctx *c, int32_t n_inst, const int32_t *inst, const int64_t *off,
                   const int64_t *vtx, double t0);
/* converge mode: per-instance, per-region scaling of both speeds (4 regions).
 * fim_set_scaling and fim_set_window only record the value on the host; the
 * values set since the last launch go to the device together before the
 * next call that runs the instance (fim_run, fim_advance, fim_run_multi) or
 * reads them (fim_residual, fim_support). */
int fim_set_scaling(fim_ctx *c, int32_t inst, const float scale[4]);
/* converge mode: the time window of one instance (us added per global
 * iteration, counted from the iteration it was seeded at; 0 = plain FIM
 * order), in place of fim_opts.window.  The fixed point does not depend on
 * it; the processing order does, and with it the work.  No counterpart in
 * the reference. */
int fim_set_window(fim_ctx *c, int32_t inst, double window);
/* converge mode: initial upper-bound field (global, nn values) with every
 * vertex active (warm start) */
int fim_init_field(fim_ctx *c, int32_t inst, const float *phi0);
/* compat mode: one reference time window -- sync_step2 then compute_step
 * (src/eikonal/cuda_fim.cu:2428-2437).  n_active (may be NULL) receives the
 * number of active vertex copies after the window. */
int fim_step(fim_ctx *c, double phi_start, double cur_time, double dt,
             double front_width, int32_t *n_active);
/* converge mode: run global iterations (interface exchange + per-submesh
 * local FIM) for all instances until no submesh is active or max_iters is
 * reached (FIM_ENOCONV).  iters (may be NULL) receives the count. */
int fim_run(fim_ctx *c, int32_t max_iters, int32_t *iters);
/* Continuous batching (no counterpart in the reference, which solves one map
 * per process).  fim_advance runs at least `iters` global iterations (whole
 * captured graphs) or until nothing is queued, without FIM_ENOCONV; *left
 * (may be NULL) receives the queued work units left.  Device counters
 * accumulate until a call with reset_stats != 0.  fim_active writes, per
 * instance, 1 if it still has queued work and 0 if its map is converged.
 * fim_recycle clears converged instances (times to +inf) and restarts their
 * time windows at the current iteration, so they can be seeded again
 * (fim_seed / fim_seed_batch) while the others keep running: each map comes
 * out bit-identical to a solve started on its own. */
int fim_advance(fim_ctx *c, int32_t iters, int32_t reset_stats, int32_t *left);
/* fim_run for n distinct converge-mode contexts at once (no counterpart in
 * the reference): one host thread keeps every context's stream fed, so the
 * contexts' kernels share the device and one context's launch tail is
 * filled by another's work (a batch split over two contexts of half the
 * instances runs faster than one context; every map is the same as from
 * fim_run).  iters (may be NULL) receives each context's global iterations;
 * FIM_ENOCONV if any context reaches max_iters with work left. */
int fim_run_multi(fim_ctx *const *cs, int32_t n, int32_t max_iters, int32_t *iters);
int fim_active(fim_ctx *c, int32_t *active);
int fim_recycle(fim_ctx *c, int32_t n, const int32_t *insts);
/* global phi (min over copies; >= 1e50 means unreached, the reference uses
 * DBL_MAX) and state (max over copies; converge mode reports FROZEN for
 * reached and FAR for unreached) of one instance; either may be NULL.
 * Host merge of src/eikonal/cuda_fim.cu:2464-2485, done on the device. */
int fim_download(fim_ctx *c, int32_t inst, double *phi, int32_t *state);
/* FIM residual sweep: max over vertices of (phi - update(phi)) / (1 + phi)
 * for one instance (converge mode), the fixed-point test */
int fim_residual(fim_ctx *c, int32_t inst, double *res);
/* Support check (no counterpart in the reference): max over reached,
 * non-seed vertices (0 < phi < inf) of (best - phi) / (1 + phi), where best is
 * the smallest candidate the vertex's elements give.  A field reached from an
 * upper bound satisfies best <= phi up to the convergence threshold; a
 * larger value means some vertex started below its fixed-point time (a warm
 * start that was not an upper bound) and the map must be recomputed. */
int fim_support(fim_ctx *c, int32_t inst, double *worst);
/* fim_support without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops.  The per-vertex candidate minima live in a
 * static device array of the library, 2^20 vertices (4 MiB) per device, so
 * meshes with more vertices are rejected (FIM_ELIMIT; use fim_support).  All
 * contexts on one device share that array: calls that overlap on one GPU
 * (different host threads, or streams without ordering) race. */
int fim_support_noalloc(fim_ctx *c, int32_t inst, double *worst);
int fim_stats(const fim_ctx *c, fim_stats_t *s);
void fim_destroy(fim_ctx *c);

/* ========================================================================
 * Surrogate warm start (no counterpart in the reference)
 *
 * A coordinate MLP: 16 inputs -- vertex xyz, three pacing-site xyz (fewer
 * sites repeat the first), four regional speed factors -- 6 hidden layers of
 * 256 with SiLU, one output (activation time / t_scale).  BF16 weights and
 * activations on MFMA with FP32 accumulation, FP32 master weights, Adam.
 * Coordinates are normalised by the bounding box of the mesh given at
 * creation.  Training maps come from a converge-mode context on the same
 * mesh (device to device); a prediction shifted up by delta initialises an
 * instance as an upper bound.
 * ===================================================================== */
typedef struct fim_nn fim_nn;

typedef struct {
  int32_t device;
  uint64_t seed;         /* weight initialisation */
  double lr;             /* Adam step (0 = 1e-3) */
  double beta1, beta2;   /* Adam moments (0 = 0.9, 0.999) */
  double adam_eps;       /* (0 = 1e-8) */
  double t_scale;        /* us per output unit (0 = 1e5) */
  int32_t max_maps;      /* dataset capacity in maps (0 = 4096) */
  int32_t max_batch;     /* largest training batch (0 = 65536, multiple of 64) */
} fim_nn_opts;

#define FIM_NN_INPUTS 16
#define FIM_NN_HIDDEN 256
#define FIM_NN_LAYERS 6
/* parameters: W0[256x16] b0[256] (W_l[256x256] b_l[256]) x 5, w_out[256], b_out */
#define FIM_NN_PARAMS 333569

int fim_nn_create(const fim_mesh *m, const fim_nn_opts *o, fim_nn **out);
void fim_nn_destroy(fim_nn *nn);
/* FP32 master parameters in the order above (FIM_NN_PARAMS values) */
int fim_nn_get_params(fim_nn *nn, float *params);
int fim_nn_set_params(fim_nn *nn, const float *params);
/* one Adam step of the FP32 masters with the given gradient (FIM_NN_PARAMS
 * values, parameter order above): the update fim_nn_train applies after
 * each backward pass -- m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
 * w -= lr (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps), then the BF16 copies
 * of the hidden-layer weights.  Lets a caller (or a test) drive the
 * optimiser with a gradient of its own. */
int fim_nn_adam_step(fim_nn *nn, const float *grad);
/* Adam's first and second moments (FIM_NN_PARAMS values each, parameter
 * order; either may be NULL) and the number of steps taken (may be NULL) */
int fim_nn_get_moments(fim_nn *nn, float *m, float *v, int64_t *steps);
/* inputs of one solve: sites (n <= 3 global vertex ids), scales[4] */
int fim_nn_features(fim_nn *nn, const int64_t *vtx, const int64_t *sites, int32_t n_sites,
                    const float *scales, int64_t n, float *x);
/* dataset: append maps of instances [inst0, inst0+n) of a converge-mode
 * context on the creation mesh; sites: n x 3 (ids, -1 pads), scales: n x 4 */
int fim_nn_add_maps(fim_nn *nn, fim_ctx *c, int32_t inst0, int32_t n, const int64_t *sites,
                    const float *scales);
int fim_nn_dataset_size(const fim_nn *nn, int32_t *maps);
/* one forward + backward on given inputs (x: b x 16, t: b targets in us),
 * b a multiple of 64; returns the mean squared error (of t / t_scale) and
 * the gradient (FIM_NN_PARAMS values); the parameters do not change */
int fim_nn_gradient(fim_nn *nn, const float *x, const float *t, int32_t b, double *loss,
                    float *grad);
/* Adam training on random (map, vertex) samples of the dataset */
int fim_nn_train(fim_nn *nn, int32_t steps, int32_t batch, uint64_t seed, double *loss);
/* prediction (us) for n inputs x (n x 16) */
int fim_nn_predict(fim_nn *nn, const float *x, int64_t n, float *t);
/* prediction for every vertex of the creation mesh for one solve (us) */
int fim_nn_predict_map(fim_nn *nn, const int64_t *sites, int32_t n_sites, const float *scales,
                       float *t);
/* the latest predicted activation time (us) of each of `count` solves over
 * a vertex sample: out[j] = max over sample[0..m) of the prediction for
 * solve j (sites: count x 3 ids, -1 pads; scales: count x 4).  Solves are
 * batched into forward passes of whole samples.  Used to give each instance
 * of a batch its own time window (fim_set_window). */
int fim_nn_predict_max(fim_nn *nn, int32_t count, const int64_t *sites, const float *scales,
                       const int64_t *sample, int32_t m, float *out);
/* warm start of instance inst of c: field = prediction + delta (us),
 * sites = 0; all_active marks every vertex active (else only the times are
 * set, and the caller seeds the sites) */
int fim_nn_warm_start(fim_nn *nn, fim_ctx *c, int32_t inst, const int64_t *sites,
                      int32_t n_sites, const float *scales, double delta, int32_t all_active);

/* ========================================================================
 * Reference-format configuration and end-to-end run
 * ===================================================================== */
typedef struct {
  int32_t verbose, method_id, solvetype, rent_en;
  double dt;          /* microseconds */
  int64_t n_steps;
} fim_options_t;

/* data/default.cfg (src/io/io_json.cpp:228-296) */
int fim_config_load_options(const char *cfg_path, fim_options_t *out);
/* plan JSON (src/io/io_json.cpp:299-318): overrides opts->dt / n_steps and
 * reports the number of stimulus events and of their vertices */
int fim_config_load_plan(const char *plan_path, fim_options_t *opts, int32_t *n_events,
                         int64_t *n_vtx);
/* the events of a plan: active, start (us), duration (us), vertex count and
 * the concatenated vertex ids, in the reference's order */
int fim_config_plan_events(const char *plan_path, int32_t *active, double *tstart, double *dur,
                           int64_t *cnt, int64_t *vtx);
/* tw::run (src/main.cpp:84-104): mesh + tags + cfg + plan -> compat solve
 * -> VTK output; *avg_s = mean step-loop seconds */
int fim_run_reference(const char *mesh_path, const char *part_path, const char *plan_path,
                      const char *out_path, int32_t num_reps, const char *cfg_path,
                      int32_t device, double *avg_s);

/* device description (replaces cuda_info::print, src/core/cuda_base.cu:72-309) */
int fim_device_info(int32_t device, char *buf, int64_t len);

#ifdef __cplusplus
}
#endif
#endif /* CUFIM_H_ */
