/*
 * fim_oracle.h -- CPU restatement of the reference FIM activation-time solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle of the HIP solver:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (cufim-nvlink_amd/) never links, loads or calls it.
 *
 * Everything is plain C99, FP64, single-threaded, built with
 * -ffp-contract=off so every arithmetic step is one IEEE operation.
 *
 * Two solvers live here:
 *   orc_compat_solve  restates the reference time-window solver
 *                     (src/eikonal/cuda_fim.cu:2386-2503 solve(),
 *                      1752-1787 sync_step2, 1790-1845 update_states,
 *                      1992-2334 compute_step, 479-542 reset_states_gpu,
 *                      632-683 update_events_gpu) on the reference submesh
 *                     decomposition (src/core/types.cpp:306-570).
 *   orc_fim_solve     an FP64 label-correcting fixed-point solver for the
 *                     same discrete local update (plain FIM to convergence);
 *                     it pins the converged maps of the GPU solver.
 *
 * Pinning status of each function (which golden vector, reference build,
 * closed form or brute force checks it) is stated at its definition in
 * fim_oracle.c.
 */
#ifndef FIM_ORACLE_H_
#define FIM_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Vertex states, same numbering as the reference enum ek_state
 * (src/core/types.h:85): far, neig, narrow, next, frozen, tail, block. */
enum { ORC_FAR = 0, ORC_NEIG = 1, ORC_NARROW = 2, ORC_NEXT = 3,
       ORC_FROZEN = 4, ORC_TAIL = 5, ORC_BLOCK = 6 };

/* Local-solver variants. */
enum { ORC_LOCAL_REF = 0,     /* reference update_per_element, as written    */
       ORC_LOCAL_EXACT = 1 }; /* exact Hopf-Lax tet/tri/edge minimum          */

/* ---- element-level building blocks ------------------------------------ */

/* insertion sort of one element's node ids (src/core/sorts.h:7-22) */
void orc_sort_nodes(int64_t *con, int n);

/* metric tensor from fibre f and speeds (vf, vs):
 * M = f f^T / vf^2 + (I - f f^T) / vs^2  (src/core/types.h:174-188) */
void orc_metric(const double f[3], double vf, double vs, double M[6]);

/* 6 squared metric edge lengths of one element, order e12 e13 e23 e14 e24
 * e34 in the element's node order (src/core/types.cpp:227-257). */
void orc_geo_elem(const double *xyz, const int64_t con[4], const double M[6],
                  double geo[6]);

/* reference triangle candidate with FP32 intermediates
 * (src/eikonal/cuda_fim.h:1256-1300) */
double orc_tri_ref(double e12, double e13, double e23, double p1, double p2);

/* reference tet candidate, formula as written (src/eikonal/cuda_fim.h:1302-1360) */
/* test hook: plant a mistake in orc_tet_ref (0 = none; see fim_oracle.c) */
void orc_plant(int which);
double orc_tet_ref(double e13, double e23, double e34, double e12, double e14,
                   double e24, double p1, double p2, double p3);

/* exact candidates: minimum over the face/edge of interpolated time plus
 * metric distance; p_i >= ORC_BIG marks an unusable vertex. */
double orc_tri_exact(double e12, double e14, double e24, double p1, double p2);
double orc_tet_exact(double e12, double e13, double e23, double e14,
                     double e24, double e34, double p1, double p2, double p3);

/* candidate for local position `t` (0..3) of an element with geometry geo[6]
 * (e12 e13 e23 e14 e24 e34 in node order) and neighbour values ph[4]
 * (ph[t] ignored; values >= ORC_BIG are unusable).  variant = ORC_LOCAL_*.
 * ORC_LOCAL_REF restates update_per_element (src/eikonal/cuda_fim.h:1362-1411)
 * after its upwind filter. */
double orc_local(int variant, const double geo[6], const double ph[4], int t);

/* relabel table used by the reference for target position t: out[0..5] are
 * geo indices of (e13 e23 e34 e12 e14 e24) and out[6..8] the positions of
 * phi1..phi3 after relabelling the target to local vertex 4. */
void orc_relabel(int t, int out[9]);

#define ORC_BIG 1.0e300

/* ---- connectivity (src/core/types.cpp:43-225) ---------------------------- */

/* sorts every element, builds n2e (cnt/dsp/con) and n2n (sorted union of
 * element node sets, self included).  Outputs are malloc'ed; release with
 * orc_free.  Returns number of vertices (max id + 1). */
int64_t orc_connectivity(int64_t ne, int64_t *e2n,
                         int32_t **n2e_cnt, int64_t **n2e_dsp, int64_t **n2e_con,
                         int32_t **n2n_cnt, int64_t **n2n_dsp, int64_t **n2n_con,
                         int64_t *n2n_total);
void orc_free(void *p);

/* reference-compatible submesh numbering (src/core/types.cpp:306-570):
 * aggregate id in first-seen tag order; local vertex ids in order of first
 * appearance over the aggregate's elements (global element order, sorted
 * node order).  e2n must already be sorted per element.
 * Outputs (caller-allocated):
 *   e2a[ne]          aggregate of each element
 *   l2g_dsp[nagg+1]  offsets into l2g
 *   l2g[sum Nn_a]    global id of each local vertex
 * Returns nagg (or -1 on error).  Pass l2g == NULL to size it first:
 * then *l2g_total is set and only e2a / l2g_dsp are written. */
int64_t orc_extract_numbering(int64_t ne, const int64_t *e2n, int64_t nn,
                              const int32_t *atags, int64_t *e2a,
                              int64_t *l2g_dsp, int64_t *l2g, int64_t *l2g_total);

/* reference front width 1.01 * max edge / 0.6 over the global mesh
 * (src/eikonal/cuda_fim.h:1536-1576). e2n sorted. */
double orc_front_width(int64_t ne, const int64_t *e2n, const double *xyz);

/* ---- solvers ------------------------------------------------------------ */

typedef struct {
  int64_t nn, ne;
  const double *xyz;        /* 3*nn */
  const int64_t *e2n;       /* 4*ne, any node order (sorted internally copy) */
  const double *fib;        /* 3*ne fibre per element, NULL -> (1,0,0) */
  const double *vel;        /* 2*ne (vf, vs) per element, NULL -> 0.6, 0.6 */
  const int32_t *atags;     /* ne partition tags (compat only) */
} orc_mesh;

typedef struct {
  int32_t nev;              /* number of stimulus events */
  const int32_t *active;    /* nev */
  const double *tstart;     /* nev, microseconds */
  const int64_t *dsp;       /* nev+1 offsets into vtx */
  const int64_t *vtx;       /* global vertex ids */
} orc_events;

/* Reference time-window solve.  dt, apd in microseconds.  front_width < 0
 * means "compute it as the reference does".  Outputs: merged global phi
 * (min over copies, ORC: DBL_MAX for unreached) and state (max over copies).
 * stats (may be NULL): [0] inner iterations summed over submeshes and steps,
 * [1] vertex-slot local solves evaluated.  Returns 0 or <0 on error. */
int orc_compat_solve(const orc_mesh *m, const orc_events *ev, double dt,
                     int64_t n_steps, double apd, double front_width,
                     int variant, double *phi_out, int32_t *state_out,
                     double *stats);

/* Same windows with timed events: event i fires in the window that holds
 * its tstart (phi = tstart), every active event fires. */
/* orc_compat_solve with the step loop on `threads` host threads (OpenMP),
 * bit-identical to it; seconds[0..2] = setup, step loop, merge (may be
 * NULL).  Events fire at t = 0 as in orc_compat_solve. */
int orc_compat_solve_par(const orc_mesh *m, const orc_events *ev, double dt, int64_t n_steps,
                         double apd, double front_width, int variant, int threads,
                         double *phi_out, int32_t *state_out, double *seconds);
int orc_compat_solve_timed(const orc_mesh *m, const orc_events *ev, double dt,
                           int64_t n_steps, double apd, double front_width,
                           int variant, double *phi_out, int32_t *state_out,
                           double *stats);

/* FP64 label-correcting fixed point of the local update (plain FIM
 * semantics: every finite neighbour is usable).  Sources get value t0[i].
 * Returns 0 on success.  stats (may be NULL): [0] heap pops, [1] local solves. */
int orc_fim_solve(const orc_mesh *m, int64_t nsrc, const int64_t *src,
                  const double *t0, int variant, double *phi_out,
                  double *stats);

/* max over vertices of (phi[v] - min(phi[v], update(v))) / (1 + phi[v]):
 * the relative amount by which any vertex can still be lowered. */
double orc_residual(const orc_mesh *m, const double *phi, int variant);

#ifdef __cplusplus
}
#endif
#endif
