/*
 * fim_oracle.c -- CPU restatement of the reference FIM solve (test oracle only).
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library; the product never does.
 *
 * Built with -O2 -ffp-contract=off (see oracle/Makefile): every arithmetic
 * step below is one IEEE-754 double (or float, where marked) operation.
 *
 * Pinning (what independent evidence checks each function; see
 * tests/test_oracle.py):
 *   orc_sort_nodes, orc_connectivity, orc_extract_numbering, orc_geo_elem,
 *   orc_metric, orc_front_width
 *       -> bit-exact against the reference host code compiled from its own
 *          sources into oracle/_ref (when built) and brute force on tiny
 *          meshes; metric against the closed form e^T M e.
 *   orc_relabel      -> the reference relabel tables (fixture) and the
 *                       Gram-matrix identity check of test_gray_code.
 *   orc_tri_ref      -> closed-form stationary point and brute-force 1-D
 *                       minimisation over lambda.
 *   orc_tet_ref      -> reproduces the reference formula as written; pinned
 *                       only through the golden replay (cuda_fim_res.vtk).
 *   orc_tri_exact, orc_tet_exact, orc_local(EXACT)
 *                    -> brute-force minimisation over the face (1-D / 2-D),
 *                       plane-wave exactness.
 *   orc_compat_solve -> golden replay of data/output/cuda_fim_res.vtk.
 *   orc_fim_solve    -> zero residual at the fixed point, point-source
 *                       lower bound (metric distance) and Dijkstra upper
 *                       bound, plane-wave exactness on aligned meshes.
 */
#include "fim_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define DMAX DBL_MAX /* reference "infinity" (src/core/types.h:30) */

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* element-level pieces                                                      */
/* ------------------------------------------------------------------------ */

/* src/core/sorts.h:7-22 (insertion sort, ascending) */
void orc_sort_nodes(int64_t *con, int n) {
  for (int i = 1; i < n; ++i) {
    int64_t key = con[i];
    int j = i - 1;
    while (j >= 0 && con[j] > key) {
      con[j + 1] = con[j];
      --j;
    }
    con[j + 1] = key;
  }
}

/* src/core/types.h:174-188 compute_velocity_tensor_f3: same operation order
 * so the reference host values are reproduced bit for bit. */
void orc_metric(const double f[3], double vfs, double vss, double M[6]) {
  const double f0 = f[0], f1 = f[1], f2 = f[2];
  const double fa = f0 * f0, fb = f0 * f1, fc = f0 * f2;
  const double fd = f1 * f1, fe = f1 * f2, ff = f2 * f2;
  const double vf = 1. / vfs / vfs, vs = 1. / vss / vss;
  M[0] = vf * fa + vs * (1. - fa);
  M[1] = (vf - vs) * fb;
  M[2] = (vf - vs) * fc;
  M[3] = vf * fd + vs * (1. - fd);
  M[4] = (vf - vs) * fe;
  M[5] = vf * ff + vs * (1. - ff);
}

/* e^T M e with the operation order of aMb_sym (src/core/types.h:165-171) */
static double quad_form(const double e[3], const double M[6]) {
  return (M[0] * e[0] + M[1] * e[1] + M[2] * e[2]) * e[0] +
         (M[1] * e[0] + M[3] * e[1] + M[4] * e[2]) * e[1] +
         (M[2] * e[0] + M[4] * e[1] + M[5] * e[2]) * e[2];
}

/* edge vector x_a - x_b (load_edge(a, b) in src/core/types.h:203-208) */
static void edge_vec(const double *xyz, int64_t a, int64_t b, double e[3]) {
  e[0] = xyz[3 * a + 0] - xyz[3 * b + 0];
  e[1] = xyz[3 * a + 1] - xyz[3 * b + 1];
  e[2] = xyz[3 * a + 2] - xyz[3 * b + 2];
}

/* src/core/types.cpp:240-253: e12 e13 e23 e14 e24 e34 */
void orc_geo_elem(const double *xyz, const int64_t con[4], const double M[6],
                  double geo[6]) {
  static const int pa[6] = {0, 0, 1, 0, 1, 2};
  static const int pb[6] = {1, 2, 2, 3, 3, 3};
  double e[3];
  for (int k = 0; k < 6; ++k) {
    edge_vec(xyz, con[pa[k]], con[pb[k]], e);
    geo[k] = quad_form(e, M);
  }
}

/* index of the edge between node positions i and j in geo[6] */
static int edge_index(int i, int j) {
  if (i > j) {
    int t = i;
    i = j;
    j = t;
  }
  /* (0,1)->0 (0,2)->1 (1,2)->2 (0,3)->3 (1,3)->4 (2,3)->5 */
  if (j == 3) return 3 + i;
  return i + j - 1;
}

/* The reference relabels an element so that the target is local vertex 4 and
 * the others follow cyclically (src/eikonal/cuda_fim.h:41-48 tables, used at
 * 1384-1392).  Derived here from the cyclic order instead of stored. */
void orc_relabel(int t, int out[9]) {
  const int q1 = (t + 1) & 3, q2 = (t + 2) & 3, q3 = (t + 3) & 3;
  out[0] = edge_index(q1, q3); /* e13 */
  out[1] = edge_index(q2, q3); /* e23 */
  out[2] = edge_index(q3, t);  /* e34 */
  out[3] = edge_index(q1, q2); /* e12 */
  out[4] = edge_index(q1, t);  /* e14 */
  out[5] = edge_index(q2, t);  /* e24 */
  out[6] = q1;
  out[7] = q2;
  out[8] = q3;
}

/* src/eikonal/cuda_fim.h:1256-1300.  The stationary points of
 *   phi2 - l*(phi2-phi1) + sqrt(a + 2 l b + l^2 c)
 * on the edge (1,2) seen from vertex 3 of the triangle; intermediates are
 * single precision exactly where the reference declares float. */
double orc_tri_ref(double e12, double e13, double e23, double phi1,
                   double phi2) {
  const double a = e23;
  const double b = -0.5 * (e12 + e23 - e13);
  const double c = e12;
  const double phi12 = phi2 - phi1;
  float dv = (float)(phi12 * phi12 - c);
  const float sv = (float)((double)dv * (b * b - a * c));
  if (dv != 0.0 && sv >= 0.0) {
    dv = (float)(c * (double)dv);
    const float sq = sqrtf(sv);
    const double base = b * phi12 * phi12 - b * c;
    const float l1 = (float)(-(base - (double)sq * phi12) / (double)dv);
    const float l2 = (float)(-(base + (double)sq * phi12) / (double)dv);
    float r1 = (float)(phi2 - (double)l1 * phi12 +
                       sqrt(a + 2.0 * (double)l1 * b + (double)(l1 * l1) * c));
    if (!((l1 >= 0.0) && (l1 <= 1.0))) r1 = INFINITY;
    float r2 = (float)(phi2 - (double)l2 * phi12 +
                       sqrt(a + 2.0 * (double)l2 * b + (double)(l2 * l2) * c));
    if (!((l2 >= 0.0) && (l2 <= 1.0))) r2 = INFINITY;
    return (double)fminf(r1, r2);
  }
  return DMAX;
}

/* src/eikonal/cuda_fim.h:1302-1360, reproduced as written (including its
 * coefficient naming, which does not match the closed form it evaluates). */
/* Planted mistakes, for tests only (orc_plant): the golden replay must
 * reject each.  1: sqrt(sqt) where the formula has sqt; 2: the coefficients
 * named as the distance term reads them (a = e34.e34, b = e13.e34,
 * c = e13.e13, d = e13.e23, e = e23.e34, f = e23.e23); 3: no tet candidate
 * at all (triangles and edges only). */
static int g_plant = 0;
void orc_plant(int which) { g_plant = which; }

double orc_tet_ref(double e13, double e23, double e34, double e12, double e14,
                   double e24, double phi1, double phi2, double phi3) {
  if (g_plant == 3) return DMAX;
  const double a0 = e13, b0 = 0.5 * (e23 + e13 - e12), c0 = -0.5 * (e34 + e13 - e14);
  const double d0 = e23, e0 = -0.5 * (e34 + e23 - e24), f0 = e34;
  const int rn = g_plant == 2;
  const double a = rn ? f0 : a0;
  const double b = rn ? c0 : b0;
  const double c = rn ? a0 : c0;
  const double d = rn ? b0 : d0;
  const double e = e0;
  const double f = rn ? d0 : f0;
  const double p13 = phi3 - phi1;
  const double p23 = phi3 - phi2;

  const double sqt =
      -e * e * c * d * d + a * d * d * d * d - (b * b * c - a * c * c) * f * f +
      ((b * b - a * c) * f * f - (e * e * c - a * d * d - 2. * (e * c - b * d) * e) * f) * p13 * p13 -
      2. * (e * e * c * d - 2. * e * b * d * d + a * d * d * d + (b * b - a * c) * d * f) * p13 * p23 +
      (e * e * c * c - 2. * e * b * c * d + a * c * d * d + (b * b * c - a * c * c) * f) * p23 * p23 +
      2. * (e * c * d * d - b * d * d * d) * e +
      (e * e * c * c + (b * b - 2. * a * c) * d * d - 2. * (e * c * c - b * c * d) * e) * f;
  if (sqt < 0.) return DMAX;
  const double sq = g_plant == 1 ? sqrt(sqt) : sqt;

  const double q1 = e * c * d * d - b * d * d * d + (e * c - b * d) * f * p13 * p13 -
                    2. * (e * c * d - b * d * d) * p13 * p23 +
                    (e * c * c - b * c * d) * p23 * p23 - (e * c * c - b * c * d) * f;
  const double q2 = (d * d * d * d - 2. * c * d * d * f + c * c * f * f +
                     (d * d * f - c * f * f) * p13 * p13 -
                     2. * (d * d * d - c * d * f) * p13 * p23 +
                     (c * d * d - c * c * f) * p23 * p23);
  const double m1 = (q1 + sq * (d * p13 - c * p23)) / q2;
  const double m2 = (q1 - sq * (d * p13 - c * p23)) / q2;
  const double k1 = -((f * m1 + e) * p13 - (d * m1 + b) * p23) / (d * p13 - c * p23);
  const double k2 = -((f * m2 + e) * p13 - (d * m2 + b) * p23) / (d * p13 - c * p23);

  double w = 1 - k1 - m1, r1 = DMAX, r2 = DMAX;
  if ((k1 > 0. && k1 < 1.) && (m1 > 0. && m1 < 1.) && w > 0.)
    r1 = k1 * phi1 + m1 * phi2 + w * phi3 +
         sqrt(a + 2. * k1 * b + k1 * k1 * c + 2. * k1 * m1 * d + 2. * m1 * e + m1 * m1 * f);
  w = 1 - k2 - m2;
  if ((k2 > 0. && k2 < 1.) && (m2 > 0. && m2 < 1.) && w > 0.)
    r2 = k2 * phi1 + m2 * phi2 + w * phi3 +
         sqrt(a + 2. * k2 * b + k2 * k2 * c + 2. * k2 * m2 * d + 2. * m2 * e + m2 * m2 * f);
  return fmin(r1, r2);
}

/* Exact minimum over the open edge (1,2) of phi(p) + |x4 - p|_M, with
 * p = x1 + l (x2 - x1).  Stationary point: l = (g - s d)/G with
 * s^2 (1 - d^2/G) = |e14|^2 - g^2/G, G = |e12|^2, g = e12.e14, d = p2 - p1. */
double orc_tri_exact(double e12, double e14, double e24, double p1, double p2) {
  if (p1 >= ORC_BIG || p2 >= ORC_BIG) return ORC_BIG;
  const double G = e12, g = 0.5 * (e12 + e14 - e24), d = p2 - p1;
  if (G <= 0.0 || d * d >= G) return ORC_BIG;
  const double r = e14 - g * g / G;
  if (r <= 0.0) return ORC_BIG;
  const double s = sqrt(r * G / (G - d * d));
  const double l = (g - s * d) / G;
  if (!(l > 0.0 && l < 1.0)) return ORC_BIG;
  return p1 + l * d + s;
}

/* Exact minimum over the open face (1,2,3) of phi(p) + |x4 - p|_M, with
 * p = x3 + l1 (x1 - x3) + l2 (x2 - x3).  With Gram matrix G of A = x1-x3,
 * B = x2-x3, g = (A.u, B.u), u = x4 - x3, d = (p1-p3, p2-p3):
 *   l = G^-1 (g - s d),  s^2 (1 - d^T G^-1 d) = |u|^2 - g^T G^-1 g. */
double orc_tet_exact(double e12, double e13, double e23, double e14,
                     double e24, double e34, double p1, double p2, double p3) {
  if (p1 >= ORC_BIG || p2 >= ORC_BIG || p3 >= ORC_BIG) return ORC_BIG;
  const double AA = e13, BB = e23, AB = 0.5 * (e13 + e23 - e12);
  const double Au = 0.5 * (e13 + e34 - e14), Bu = 0.5 * (e23 + e34 - e24);
  const double d1 = p1 - p3, d2 = p2 - p3;
  const double det = AA * BB - AB * AB;
  if (det <= 0.0) return ORC_BIG;
  const double w = (BB * d1 * d1 - 2.0 * AB * d1 * d2 + AA * d2 * d2) / det;
  if (w >= 1.0) return ORC_BIG;
  const double r = e34 - (BB * Au * Au - 2.0 * AB * Au * Bu + AA * Bu * Bu) / det;
  if (r <= 0.0) return ORC_BIG;
  const double s = sqrt(r / (1.0 - w));
  const double t1 = Au - s * d1, t2 = Bu - s * d2;
  const double l1 = (BB * t1 - AB * t2) / det;
  const double l2 = (AA * t2 - AB * t1) / det;
  if (!(l1 > 0.0 && l2 > 0.0 && l1 + l2 < 1.0)) return ORC_BIG;
  return p3 + l1 * d1 + l2 * d2 + s;
}

double orc_local(int variant, const double geo[6], const double ph[4], int t) {
  int r[9];
  orc_relabel(t, r);
  const double e13 = geo[r[0]], e23 = geo[r[1]], e34 = geo[r[2]];
  const double e12 = geo[r[3]], e14 = geo[r[4]], e24 = geo[r[5]];
  const double p1 = ph[r[6]], p2 = ph[r[7]], p3 = ph[r[8]];
  if (variant == ORC_LOCAL_REF) {
    /* src/eikonal/cuda_fim.h:1395-1405 */
    double v = orc_tet_ref(e13, e23, e34, e12, e14, e24, p1, p2, p3);
    v = fmin(v, orc_tri_ref(e12, e14, e24, p1, p2));
    v = fmin(v, orc_tri_ref(e23, e24, e34, p2, p3));
    v = fmin(v, orc_tri_ref(e13, e14, e34, p1, p3));
    v = fmin(v, p1 + sqrt(e14));
    v = fmin(v, p2 + sqrt(e24));
    v = fmin(v, p3 + sqrt(e34));
    return v;
  }
  double v = ORC_BIG;
  if (p1 < ORC_BIG) v = fmin(v, p1 + sqrt(e14));
  if (p2 < ORC_BIG) v = fmin(v, p2 + sqrt(e24));
  if (p3 < ORC_BIG) v = fmin(v, p3 + sqrt(e34));
  v = fmin(v, orc_tri_exact(e12, e14, e24, p1, p2));
  v = fmin(v, orc_tri_exact(e13, e14, e34, p1, p3));
  v = fmin(v, orc_tri_exact(e23, e24, e34, p2, p3));
  v = fmin(v, orc_tet_exact(e12, e13, e23, e14, e24, e34, p1, p2, p3));
  return v >= ORC_BIG ? DMAX : v; /* no candidate: the reference sentinel */
}

/* ------------------------------------------------------------------------ */
/* connectivity                                                              */
/* ------------------------------------------------------------------------ */

static int cmp_i64(const void *a, const void *b) {
  const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

/* n2e: vertex -> incident elements in ascending element order
 * (src/core/types.cpp:83-129). */
static void build_n2e(int64_t ne, const int64_t *e2n, int64_t nn, int32_t *cnt,
                      int64_t *dsp, int64_t *con) {
  memset(cnt, 0, sizeof(int32_t) * nn);
  for (int64_t i = 0; i < 4 * ne; ++i) cnt[e2n[i]]++;
  int64_t acc = 0;
  for (int64_t v = 0; v < nn; ++v) {
    dsp[v] = acc;
    acc += cnt[v];
  }
  memset(cnt, 0, sizeof(int32_t) * nn);
  for (int64_t e = 0; e < ne; ++e)
    for (int k = 0; k < 4; ++k) {
      const int64_t v = e2n[4 * e + k];
      con[dsp[v] + cnt[v]++] = e;
    }
}

/* n2n: sorted union of the node sets of the incident elements, the vertex
 * itself included (src/core/types.cpp:139-224). */
static int64_t build_n2n(int64_t ne, const int64_t *e2n, int64_t nn,
                         const int32_t *n2e_cnt, const int64_t *n2e_dsp,
                         const int64_t *n2e_con, int32_t *cnt, int64_t *dsp,
                         int64_t **con_out) {
  int64_t cap = 0;
  for (int64_t v = 0; v < nn; ++v) cap += 4 * (int64_t)n2e_cnt[v];
  int64_t *con = (int64_t *)malloc(sizeof(int64_t) * (cap > 0 ? cap : 1));
  int64_t *buf = (int64_t *)malloc(sizeof(int64_t) * 4096);
  int64_t acc = 0;
  (void)ne;
  for (int64_t v = 0; v < nn; ++v) {
    int n = 0;
    for (int32_t k = 0; k < n2e_cnt[v]; ++k) {
      const int64_t e = n2e_con[n2e_dsp[v] + k];
      for (int j = 0; j < 4; ++j) buf[n++] = e2n[4 * e + j];
    }
    qsort(buf, n, sizeof(int64_t), cmp_i64);
    int u = 0;
    for (int i = 0; i < n; ++i)
      if (u == 0 || buf[i] != buf[u - 1]) buf[u++] = buf[i];
    dsp[v] = acc;
    cnt[v] = u;
    memcpy(con + acc, buf, sizeof(int64_t) * u);
    acc += u;
  }
  free(buf);
  *con_out = con;
  return acc;
}

int64_t orc_connectivity(int64_t ne, int64_t *e2n, int32_t **n2e_cnt,
                         int64_t **n2e_dsp, int64_t **n2e_con, int32_t **n2n_cnt,
                         int64_t **n2n_dsp, int64_t **n2n_con,
                         int64_t *n2n_total) {
  int64_t nn = 0;
  for (int64_t e = 0; e < ne; ++e) orc_sort_nodes(e2n + 4 * e, 4);
  for (int64_t i = 0; i < 4 * ne; ++i)
    if (e2n[i] + 1 > nn) nn = e2n[i] + 1;
  *n2e_cnt = (int32_t *)malloc(sizeof(int32_t) * nn);
  *n2e_dsp = (int64_t *)malloc(sizeof(int64_t) * nn);
  *n2e_con = (int64_t *)malloc(sizeof(int64_t) * 4 * ne);
  build_n2e(ne, e2n, nn, *n2e_cnt, *n2e_dsp, *n2e_con);
  *n2n_cnt = (int32_t *)malloc(sizeof(int32_t) * nn);
  *n2n_dsp = (int64_t *)malloc(sizeof(int64_t) * nn);
  *n2n_total = build_n2n(ne, e2n, nn, *n2e_cnt, *n2e_dsp, *n2e_con, *n2n_cnt,
                         *n2n_dsp, n2n_con);
  return nn;
}

/* src/core/types.cpp:329-408: aggregate ids in first-seen tag order, local
 * vertex ids in order of first appearance over each aggregate's elements. */
int64_t orc_extract_numbering(int64_t ne, const int64_t *e2n, int64_t nn,
                              const int32_t *atags, int64_t *e2a,
                              int64_t *l2g_dsp, int64_t *l2g,
                              int64_t *l2g_total) {
  /* tag -> aggregate id (first seen) via sorted unique tags */
  int64_t nagg = 0;
  int32_t *seen_tag = (int32_t *)malloc(sizeof(int32_t) * (ne > 0 ? ne : 1));
  for (int64_t e = 0; e < ne; ++e) {
    int64_t a = -1;
    for (int64_t k = 0; k < nagg; ++k)
      if (seen_tag[k] == atags[e]) {
        a = k;
        break;
      }
    if (a < 0) {
      seen_tag[nagg] = atags[e];
      a = nagg++;
    }
    e2a[e] = a;
  }
  free(seen_tag);
  /* group elements per aggregate, keeping global order */
  int64_t *a_cnt = (int64_t *)calloc(nagg + 1, sizeof(int64_t));
  for (int64_t e = 0; e < ne; ++e) a_cnt[e2a[e] + 1]++;
  for (int64_t a = 0; a < nagg; ++a) a_cnt[a + 1] += a_cnt[a];
  int64_t *a2e = (int64_t *)malloc(sizeof(int64_t) * (ne > 0 ? ne : 1));
  int64_t *fill = (int64_t *)calloc(nagg > 0 ? nagg : 1, sizeof(int64_t));
  for (int64_t e = 0; e < ne; ++e) a2e[a_cnt[e2a[e]] + fill[e2a[e]]++] = e;
  free(fill);
  int64_t *stamp = (int64_t *)malloc(sizeof(int64_t) * (nn > 0 ? nn : 1));
  for (int64_t v = 0; v < nn; ++v) stamp[v] = -1;
  int64_t total = 0;
  for (int64_t a = 0; a < nagg; ++a) {
    l2g_dsp[a] = total;
    int64_t cnt = 0;
    for (int64_t i = a_cnt[a]; i < a_cnt[a + 1]; ++i) {
      const int64_t e = a2e[i];
      for (int k = 0; k < 4; ++k) {
        const int64_t v = e2n[4 * e + k];
        if (stamp[v] != a) {
          stamp[v] = a;
          if (l2g) l2g[total + cnt] = v;
          cnt++;
        }
      }
    }
    total += cnt;
  }
  l2g_dsp[nagg] = total;
  if (l2g_total) *l2g_total = total;
  free(stamp);
  free(a2e);
  free(a_cnt);
  return nagg;
}

/* src/eikonal/cuda_fim.h:1536-1576 */
double orc_front_width(int64_t ne, const int64_t *e2n, const double *xyz) {
  /* the reference maximises squared lengths over n2n pairs, i.e. over all
   * vertex pairs that share an element */
  double best = 0.0;
  for (int64_t e = 0; e < ne; ++e)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        const double *A = xyz + 3 * e2n[4 * e + i], *B = xyz + 3 * e2n[4 * e + j];
        const double x = B[0] - A[0], y = B[1] - A[1], z = B[2] - A[2];
        const double t = x * x + y * y + z * z;
        if (t > best) best = t;
      }
  return 1.01 * sqrt(best) / 0.6;
}

/* ------------------------------------------------------------------------ */
/* compat (reference time-window) solver                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
  int64_t nn, ne, nenc;
  int64_t *e2n; /* local, sorted per element */
  int32_t *n2e_cnt;
  int64_t *n2e_dsp, *n2e_con;
  int32_t *n2n_cnt;
  int64_t *n2n_dsp, *n2n_con;
  double *geo;
  int64_t *l2g;
  int8_t *tpos; /* position of the slot owner inside the slot's element */
  double *phi, *tmp_phi;
  int32_t *state, *tmp_state, *snap;
} sub_t;

static void sub_free(sub_t *s) {
  free(s->e2n); free(s->n2e_cnt); free(s->n2e_dsp); free(s->n2e_con);
  free(s->n2n_cnt); free(s->n2n_dsp); free(s->n2n_con); free(s->geo);
  free(s->l2g); free(s->tpos); free(s->phi); free(s->tmp_phi);
  free(s->state); free(s->tmp_state); free(s->snap);
}

static void elem_metric(const orc_mesh *m, int64_t e, double M[6]) {
  static const double f0[3] = {1.0, 0.0, 0.0};
  const double *f = m->fib ? m->fib + 3 * e : f0;
  const double vf = m->vel ? m->vel[2 * e] : 0.6;
  const double vs = m->vel ? m->vel[2 * e + 1] : 0.6;
  orc_metric(f, vf, vs, M);
}

/* slot candidate, src/eikonal/cuda_fim.h:1362-1411 with its upwind filter
 * (state >= narrow and phi > phi_start - front_width). */
static double slot_update(const sub_t *s, int64_t slot, int64_t owner,
                          int variant, double phi_start, double fw) {
  const int64_t e = s->n2e_con[slot];
  const int64_t *con = s->e2n + 4 * e;
  double ph[4];
  const int t = s->tpos[slot];
  for (int k = 0; k < 4; ++k) {
    const int64_t n = con[k];
    if (n == owner) {
      ph[k] = DMAX;
      continue;
    }
    ph[k] = (s->state[n] >= ORC_NARROW && s->phi[n] > phi_start - fw)
                ? s->phi[n] : DMAX;
  }
  return orc_local(variant, s->geo + 6 * e, ph, t);
}

/* src/eikonal/cuda_fim.cu:1790-1845 */
static int64_t update_states(sub_t *s, double cur_time, double apd) {
  int64_t active = 0;
  for (int64_t v = 0; v < s->nn; ++v) {
    const int32_t st = s->state[v];
    if (st == ORC_NARROW) {
      active++;
    } else if (st == ORC_NEXT) {
      active++;
      s->state[v] = ORC_NARROW;
    } else if (st == ORC_FROZEN) {
      if (!(s->phi[v] + apd > cur_time)) s->state[v] = ORC_TAIL;
    } else if (st == ORC_TAIL) {
      if (!(s->phi[v] + apd / 0.91 > cur_time)) {
        s->phi[v] = DMAX;
        s->state[v] = ORC_FAR;
      }
    }
  }
  return active;
}

/* src/eikonal/cuda_fim.cu:1992-2334 (phases a-g); phase c reads the
 * just-frozen flags as they stood after phase b. */
static void compute_step(sub_t *s, int variant, double fw, double cur_time,
                         double phi_start, double dt, double apd, double *stats) {
  int64_t active = update_states(s, cur_time, apd);
  while (active > 0) {
    active = 0;
    if (stats) stats[0] += 1;
    /* a: slot candidates of narrow owners */
    for (int64_t v = 0; v < s->nn; ++v)
      for (int32_t k = 0; k < s->n2e_cnt[v]; ++k) {
        const int64_t slot = s->n2e_dsp[v] + k;
        s->tmp_phi[slot] = DMAX;
        if (s->state[v] == ORC_NARROW) {
          s->tmp_phi[slot] = slot_update(s, slot, v, variant, phi_start, fw);
          if (stats) stats[1] += 1;
        }
      }
    /* b: reduce, window test, convergence test */
    for (int64_t v = 0; v < s->nn; ++v) {
      s->tmp_state[v] = ORC_FAR;
      if (s->state[v] != ORC_NARROW) continue;
      const double old = s->phi[v];
      double nw = old;
      for (int32_t k = 0; k < s->n2e_cnt[v]; ++k) {
        const double c = s->tmp_phi[s->n2e_dsp[v] + k];
        if (c < nw) nw = c;
      }
      s->phi[v] = nw;
      if (nw - phi_start > dt)
        s->state[v] = ORC_NEXT;
      else if (fabs(old - nw) < (1e-6 * (1 + (old + nw) * .5)))
        s->tmp_state[v] = ORC_FROZEN;
    }
    /* c: mark downwind neighbours of just-frozen vertices */
    memcpy(s->snap, s->tmp_state, sizeof(int32_t) * s->nn);
    for (int64_t v = 0; v < s->nn; ++v) {
      if (s->state[v] == ORC_NARROW) continue;
      for (int32_t k = 0; k < s->n2n_cnt[v]; ++k) {
        const int64_t n = s->n2n_con[s->n2n_dsp[v] + k];
        if (s->snap[n] == ORC_FROZEN && s->phi[n] < s->phi[v]) {
          s->tmp_state[v] = s->state[v];
          s->state[v] = ORC_NEIG;
          break;
        }
      }
    }
    /* d: slot candidates of marked vertices */
    for (int64_t v = 0; v < s->nn; ++v)
      for (int32_t k = 0; k < s->n2e_cnt[v]; ++k) {
        const int64_t slot = s->n2e_dsp[v] + k;
        s->tmp_phi[slot] = DMAX;
        if (s->state[v] == ORC_NEIG) {
          s->tmp_phi[slot] = slot_update(s, slot, v, variant, phi_start, fw);
          if (stats) stats[1] += 1;
        }
      }
    /* e: accept improvements */
    for (int64_t v = 0; v < s->nn; ++v) {
      if (s->state[v] != ORC_NEIG) continue;
      const double old = s->phi[v];
      double nw = old;
      for (int32_t k = 0; k < s->n2e_cnt[v]; ++k) {
        const double c = s->tmp_phi[s->n2e_dsp[v] + k];
        if (c < nw) nw = c;
      }
      s->phi[v] = nw;
      s->state[v] = (nw < old) ? ORC_NARROW : s->tmp_state[v];
    }
    /* f: freeze / count */
    for (int64_t v = 0; v < s->nn; ++v) {
      if (s->state[v] == ORC_NARROW) {
        if (s->tmp_state[v] == ORC_FROZEN)
          s->state[v] = ORC_FROZEN;
        else
          active++;
      } else if (s->state[v] == ORC_NEIG) {
        s->state[v] = ORC_NARROW;
        active++;
      }
    }
  }
}

static int compat_solve(const orc_mesh *m, const orc_events *ev, double dt, int64_t n_steps,
                        double apd, double front_width, int variant, double *phi_out,
                        int32_t *state_out, double *stats, int timed, int threads,
                        double *seconds);

int orc_compat_solve(const orc_mesh *m, const orc_events *ev, double dt,
                     int64_t n_steps, double apd, double front_width,
                     int variant, double *phi_out, int32_t *state_out,
                     double *stats) {
  return compat_solve(m, ev, dt, n_steps, apd, front_width, variant, phi_out, state_out, stats, 0,
                      1, NULL);
}

/* The same solve with the step loop on `threads` host threads (OpenMP):
 * the interface merge over global vertices and compute_step over submeshes,
 * one submesh per thread at a time, as the reference runs one block per
 * submesh.  Submeshes share nothing within a compute phase, so the result
 * is bit-identical to orc_compat_solve.  seconds[0..2] (may be NULL)
 * receive the setup, step-loop and merge times; the step loop alone is what
 * the reference times (cuda_fim.cu:2424-2447). */
int orc_compat_solve_par(const orc_mesh *m, const orc_events *ev, double dt, int64_t n_steps,
                         double apd, double front_width, int variant, int threads,
                         double *phi_out, int32_t *state_out, double *seconds) {
  return compat_solve(m, ev, dt, n_steps, apd, front_width, variant, phi_out, state_out, NULL, 0,
                      threads > 0 ? threads : 1, seconds);
}

/* Timed events: the same windows, but event i fires in the window
 * [phi_start, cur) that contains its tstart, seeding its vertices with
 * phi = tstart and state narrow before that window's sync; every active
 * event fires (no stop at the first miss).  This is the multi-beat
 * behaviour the reference's update_events_gpu (cuda_fim.cu:632-683) would
 * give if it compared tstart with the window rather than firing everything
 * at cur_time 0; it has no golden and is pinned by the host-side checks in
 * tests/test_oracle.py and the GPU parity test. */
int orc_compat_solve_timed(const orc_mesh *m, const orc_events *ev, double dt, int64_t n_steps,
                           double apd, double front_width, int variant, double *phi_out,
                           int32_t *state_out, double *stats) {
  return compat_solve(m, ev, dt, n_steps, apd, front_width, variant, phi_out, state_out, stats, 1,
                      1, NULL);
}

static int compat_solve(const orc_mesh *m, const orc_events *ev, double dt, int64_t n_steps,
                        double apd, double front_width, int variant, double *phi_out,
                        int32_t *state_out, double *stats, int timed, int threads,
                        double *seconds) {
  const double t_setup0 = omp_get_wtime();
  const int64_t ne = m->ne, nn = m->nn;
  if (ne <= 0 || nn <= 0 || !m->atags) return -1;
  int64_t *e2n = (int64_t *)malloc(sizeof(int64_t) * 4 * ne);
  memcpy(e2n, m->e2n, sizeof(int64_t) * 4 * ne);
  for (int64_t e = 0; e < ne; ++e) orc_sort_nodes(e2n + 4 * e, 4);
  for (int64_t i = 0; i < 4 * ne; ++i)
    if (e2n[i] < 0 || e2n[i] >= nn) { free(e2n); return -2; }

  const double fw = front_width >= 0.0 ? front_width : orc_front_width(ne, e2n, m->xyz);
  int64_t *e2a = (int64_t *)malloc(sizeof(int64_t) * ne);
  int64_t *l2g_dsp = (int64_t *)malloc(sizeof(int64_t) * (ne + 1));
  int64_t total = 0;
  const int64_t nagg = orc_extract_numbering(ne, e2n, nn, m->atags, e2a, l2g_dsp, NULL, &total);
  int64_t *l2g = (int64_t *)malloc(sizeof(int64_t) * (total > 0 ? total : 1));
  orc_extract_numbering(ne, e2n, nn, m->atags, e2a, l2g_dsp, l2g, &total);

  sub_t *subs = (sub_t *)calloc(nagg, sizeof(sub_t));
  int64_t *g2l = (int64_t *)malloc(sizeof(int64_t) * nn);
  for (int64_t v = 0; v < nn; ++v) g2l[v] = -1;
  /* elements of each aggregate in ascending order (a counting sort by e2a) */
  int64_t *a_dsp = (int64_t *)calloc(nagg + 1, sizeof(int64_t));
  int64_t *a_el = (int64_t *)malloc(sizeof(int64_t) * ne);
  for (int64_t e = 0; e < ne; ++e) a_dsp[e2a[e] + 1]++;
  for (int64_t a = 0; a < nagg; ++a) a_dsp[a + 1] += a_dsp[a];
  {
    int64_t *fill = (int64_t *)calloc(nagg, sizeof(int64_t));
    for (int64_t e = 0; e < ne; ++e) a_el[a_dsp[e2a[e]] + fill[e2a[e]]++] = e;
    free(fill);
  }
  for (int64_t a = 0; a < nagg; ++a) {
    sub_t *s = subs + a;
    s->nn = l2g_dsp[a + 1] - l2g_dsp[a];
    s->l2g = (int64_t *)malloc(sizeof(int64_t) * s->nn);
    memcpy(s->l2g, l2g + l2g_dsp[a], sizeof(int64_t) * s->nn);
    for (int64_t l = 0; l < s->nn; ++l) g2l[s->l2g[l]] = l;
    const int64_t cnt = a_dsp[a + 1] - a_dsp[a];
    s->ne = cnt;
    s->nenc = 4 * cnt;
    s->e2n = (int64_t *)malloc(sizeof(int64_t) * 4 * cnt);
    s->geo = (double *)malloc(sizeof(double) * 6 * cnt);
    int64_t le = 0;
    for (int64_t q = a_dsp[a]; q < a_dsp[a + 1]; ++q) {
      const int64_t e = a_el[q];
      for (int k = 0; k < 4; ++k) s->e2n[4 * le + k] = g2l[e2n[4 * e + k]];
      orc_sort_nodes(s->e2n + 4 * le, 4); /* submesh compute_base_connectivity */
      int64_t gcon[4];
      for (int k = 0; k < 4; ++k) gcon[k] = s->l2g[s->e2n[4 * le + k]];
      double M[6];
      elem_metric(m, e, M);
      orc_geo_elem(m->xyz, gcon, M, s->geo + 6 * le);
      le++;
    }
    s->n2e_cnt = (int32_t *)malloc(sizeof(int32_t) * s->nn);
    s->n2e_dsp = (int64_t *)malloc(sizeof(int64_t) * s->nn);
    s->n2e_con = (int64_t *)malloc(sizeof(int64_t) * s->nenc);
    build_n2e(s->ne, s->e2n, s->nn, s->n2e_cnt, s->n2e_dsp, s->n2e_con);
    s->n2n_cnt = (int32_t *)malloc(sizeof(int32_t) * s->nn);
    s->n2n_dsp = (int64_t *)malloc(sizeof(int64_t) * s->nn);
    build_n2n(s->ne, s->e2n, s->nn, s->n2e_cnt, s->n2e_dsp, s->n2e_con,
              s->n2n_cnt, s->n2n_dsp, &s->n2n_con);
    s->tpos = (int8_t *)malloc(s->nenc);
    for (int64_t v = 0; v < s->nn; ++v)
      for (int32_t k = 0; k < s->n2e_cnt[v]; ++k) {
        const int64_t slot = s->n2e_dsp[v] + k, e = s->n2e_con[slot];
        int t = 0;
        while (s->e2n[4 * e + t] != v) ++t;
        s->tpos[slot] = (int8_t)t;
      }
    s->phi = (double *)malloc(sizeof(double) * s->nn);
    s->tmp_phi = (double *)malloc(sizeof(double) * s->nenc);
    s->state = (int32_t *)malloc(sizeof(int32_t) * s->nn);
    s->tmp_state = (int32_t *)malloc(sizeof(int32_t) * s->nn);
    s->snap = (int32_t *)malloc(sizeof(int32_t) * s->nn);
    for (int64_t l = 0; l < s->nn; ++l) g2l[s->l2g[l]] = -1;
  }
  free(a_dsp);
  free(a_el);

  /* n2i: copies of each global vertex, ascending aggregate (types.cpp:538-545) */
  int64_t *ci_cnt = (int64_t *)calloc(nn + 1, sizeof(int64_t));
  for (int64_t a = 0; a < nagg; ++a)
    for (int64_t l = 0; l < subs[a].nn; ++l) ci_cnt[subs[a].l2g[l] + 1]++;
  for (int64_t v = 0; v < nn; ++v) ci_cnt[v + 1] += ci_cnt[v];
  int64_t *ci_a = (int64_t *)malloc(sizeof(int64_t) * (total > 0 ? total : 1));
  int64_t *ci_l = (int64_t *)malloc(sizeof(int64_t) * (total > 0 ? total : 1));
  int64_t *ci_fill = (int64_t *)calloc(nn, sizeof(int64_t));
  for (int64_t a = 0; a < nagg; ++a)
    for (int64_t l = 0; l < subs[a].nn; ++l) {
      const int64_t v = subs[a].l2g[l];
      const int64_t p = ci_cnt[v] + ci_fill[v]++;
      ci_a[p] = a;
      ci_l[p] = l;
    }

  /* reset_states_gpu (cuda_fim.cu:479-542) */
  for (int64_t a = 0; a < nagg; ++a)
    for (int64_t l = 0; l < subs[a].nn; ++l) {
      subs[a].phi[l] = DMAX;
      subs[a].state[l] = ORC_FAR;
    }
  /* update_events_gpu at cur_time 0 (cuda_fim.cu:632-683) */
  if (ev && !timed)
    for (int32_t i = 0; i < ev->nev; ++i) {
      if (!(ev->tstart[i] >= 0.0 && ev->active[i])) break;
      for (int64_t j = ev->dsp[i]; j < ev->dsp[i + 1]; ++j) {
        const int64_t v = ev->vtx[j];
        if (v < 0 || v >= nn) continue;
        for (int64_t p = ci_cnt[v]; p < ci_cnt[v + 1]; ++p) {
          subs[ci_a[p]].phi[ci_l[p]] = 0.0;
          subs[ci_a[p]].state[ci_l[p]] = ORC_NARROW;
        }
      }
    }
  if (stats) stats[0] = stats[1] = 0.0;
  const double t_loop0 = omp_get_wtime();
  double cur = 0.0;
  for (int64_t step = 0; step < n_steps; ++step) {
    const double phi_start = cur;
    cur += dt;
    if (ev && timed)
      for (int32_t i = 0; i < ev->nev; ++i) {
        if (!ev->active[i] || !(ev->tstart[i] >= phi_start && ev->tstart[i] < cur)) continue;
        for (int64_t j = ev->dsp[i]; j < ev->dsp[i + 1]; ++j) {
          const int64_t v = ev->vtx[j];
          if (v < 0 || v >= nn) continue;
          for (int64_t p = ci_cnt[v]; p < ci_cnt[v + 1]; ++p) {
            subs[ci_a[p]].phi[ci_l[p]] = ev->tstart[i];
            subs[ci_a[p]].state[ci_l[p]] = ORC_NARROW;
          }
        }
      }
    /* sync_step2 (cuda_fim.cu:1752-1787), per global vertex; a vertex's
     * copies are its own, so vertices run in parallel */
#pragma omp parallel for num_threads(threads) schedule(static, 4096) if (threads > 1)
    for (int64_t v = 0; v < nn; ++v) {
      double mn = DMAX;
      for (int64_t p = ci_cnt[v]; p < ci_cnt[v + 1]; ++p) {
        const double c = subs[ci_a[p]].phi[ci_l[p]];
        if (c < mn) mn = c;
      }
      for (int64_t p = ci_cnt[v]; p < ci_cnt[v + 1]; ++p) {
        sub_t *s = subs + ci_a[p];
        if (mn < s->phi[ci_l[p]]) {
          s->phi[ci_l[p]] = mn;
          s->state[ci_l[p]] = ORC_NARROW;
        }
      }
    }
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) if (threads > 1)
    for (int64_t a = 0; a < nagg; ++a)
      compute_step(subs + a, variant, fw, cur, phi_start, dt, apd, stats);
  }
  const double t_loop1 = omp_get_wtime();
  /* host merge (cuda_fim.cu:2464-2485) */
  for (int64_t v = 0; v < nn; ++v) {
    phi_out[v] = DMAX;
    state_out[v] = ORC_FAR;
  }
  for (int64_t a = 0; a < nagg; ++a)
    for (int64_t l = 0; l < subs[a].nn; ++l) {
      const int64_t g = subs[a].l2g[l];
      if (subs[a].phi[l] < phi_out[g]) phi_out[g] = subs[a].phi[l];
      if (subs[a].state[l] > state_out[g]) state_out[g] = subs[a].state[l];
    }

  if (seconds) {
    seconds[0] = t_loop0 - t_setup0;
    seconds[1] = t_loop1 - t_loop0;
    seconds[2] = omp_get_wtime() - t_loop1;
  }
  for (int64_t a = 0; a < nagg; ++a) sub_free(subs + a);
  free(subs); free(g2l); free(ci_cnt); free(ci_a); free(ci_l); free(ci_fill);
  free(l2g); free(l2g_dsp); free(e2a); free(e2n);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 fixed-point solver (label correcting, binary heap)                   */
/* ------------------------------------------------------------------------ */

typedef struct { double k; int64_t v; } hitem;
typedef struct { hitem *a; int64_t n, cap; } heap_t;

static void heap_push(heap_t *h, double k, int64_t v) {
  if (h->n == h->cap) {
    h->cap = h->cap ? 2 * h->cap : 1024;
    h->a = (hitem *)realloc(h->a, sizeof(hitem) * h->cap);
  }
  int64_t i = h->n++;
  while (i > 0) {
    const int64_t p = (i - 1) / 2;
    if (h->a[p].k < k || (h->a[p].k == k && h->a[p].v <= v)) break;
    h->a[i] = h->a[p];
    i = p;
  }
  h->a[i].k = k;
  h->a[i].v = v;
}

static hitem heap_pop(heap_t *h) {
  hitem top = h->a[0];
  const hitem last = h->a[--h->n];
  int64_t i = 0;
  for (;;) {
    int64_t c = 2 * i + 1;
    if (c >= h->n) break;
    if (c + 1 < h->n && (h->a[c + 1].k < h->a[c].k ||
                         (h->a[c + 1].k == h->a[c].k && h->a[c + 1].v < h->a[c].v)))
      c++;
    if (last.k < h->a[c].k || (last.k == h->a[c].k && last.v <= h->a[c].v)) break;
    h->a[i] = h->a[c];
    i = c;
  }
  if (h->n > 0) h->a[i] = last;
  return top;
}

typedef struct {
  int64_t nn, ne;
  int64_t *e2n;
  double *geo;
  int32_t *n2e_cnt;
  int64_t *n2e_dsp, *n2e_con;
} gmesh_t;

static int gmesh_build(const orc_mesh *m, gmesh_t *g) {
  g->nn = m->nn;
  g->ne = m->ne;
  g->e2n = (int64_t *)malloc(sizeof(int64_t) * 4 * m->ne);
  memcpy(g->e2n, m->e2n, sizeof(int64_t) * 4 * m->ne);
  for (int64_t i = 0; i < 4 * m->ne; ++i)
    if (g->e2n[i] < 0 || g->e2n[i] >= m->nn) return -2;
  g->geo = (double *)malloc(sizeof(double) * 6 * m->ne);
  for (int64_t e = 0; e < m->ne; ++e) {
    orc_sort_nodes(g->e2n + 4 * e, 4);
    double M[6];
    elem_metric(m, e, M);
    orc_geo_elem(m->xyz, g->e2n + 4 * e, M, g->geo + 6 * e);
  }
  g->n2e_cnt = (int32_t *)malloc(sizeof(int32_t) * m->nn);
  g->n2e_dsp = (int64_t *)malloc(sizeof(int64_t) * m->nn);
  g->n2e_con = (int64_t *)malloc(sizeof(int64_t) * 4 * m->ne);
  build_n2e(m->ne, g->e2n, m->nn, g->n2e_cnt, g->n2e_dsp, g->n2e_con);
  return 0;
}

static void gmesh_free(gmesh_t *g) {
  free(g->e2n); free(g->geo); free(g->n2e_cnt); free(g->n2e_dsp); free(g->n2e_con);
}

int orc_fim_solve(const orc_mesh *m, int64_t nsrc, const int64_t *src,
                  const double *t0, int variant, double *phi_out,
                  double *stats) {
  gmesh_t g;
  if (m->ne <= 0 || m->nn <= 0) return -1;
  if (gmesh_build(m, &g) != 0) { gmesh_free(&g); return -2; }
  const double unusable = variant == ORC_LOCAL_REF ? DMAX : ORC_BIG;
  for (int64_t v = 0; v < m->nn; ++v) phi_out[v] = DMAX;
  heap_t h = {0, 0, 0};
  for (int64_t i = 0; i < nsrc; ++i) {
    if (src[i] < 0 || src[i] >= m->nn) continue;
    const double t = t0 ? t0[i] : 0.0;
    if (t < phi_out[src[i]]) {
      phi_out[src[i]] = t;
      heap_push(&h, t, src[i]);
    }
  }
  double pops = 0, solves = 0;
  while (h.n > 0) {
    const hitem it = heap_pop(&h);
    if (it.k > phi_out[it.v]) continue;
    pops += 1;
    const int64_t v = it.v;
    for (int32_t k = 0; k < g.n2e_cnt[v]; ++k) {
      const int64_t e = g.n2e_con[g.n2e_dsp[v] + k];
      const int64_t *con = g.e2n + 4 * e;
      double ph[4];
      for (int j = 0; j < 4; ++j) ph[j] = phi_out[con[j]] < DMAX ? phi_out[con[j]] : unusable;
      for (int t = 0; t < 4; ++t) {
        const int64_t u = con[t];
        if (u == v) continue;
        const double c = orc_local(variant, g.geo + 6 * e, ph, t);
        solves += 1;
        if (c < phi_out[u] && (phi_out[u] >= DMAX || phi_out[u] - c > 1e-13 * (1.0 + c))) {
          phi_out[u] = c;
          ph[t] = c;
          heap_push(&h, c, u);
        }
      }
    }
  }
  free(h.a);
  gmesh_free(&g);
  if (stats) { stats[0] = pops; stats[1] = solves; }
  return 0;
}

double orc_residual(const orc_mesh *m, const double *phi, int variant) {
  gmesh_t g;
  if (gmesh_build(m, &g) != 0) { gmesh_free(&g); return INFINITY; }
  const double unusable = variant == ORC_LOCAL_REF ? DMAX : ORC_BIG;
  double worst = 0.0;
  for (int64_t e = 0; e < g.ne; ++e) {
    const int64_t *con = g.e2n + 4 * e;
    double ph[4];
    for (int j = 0; j < 4; ++j) ph[j] = phi[con[j]] < DMAX ? phi[con[j]] : unusable;
    for (int t = 0; t < 4; ++t) {
      const double c = orc_local(variant, g.geo + 6 * e, ph, t);
      const double cur = phi[con[t]];
      if (c < cur) {
        const double r = (cur >= DMAX) ? INFINITY : (cur - c) / (1.0 + cur);
        if (r > worst) worst = r;
      }
    }
  }
  gmesh_free(&g);
  return worst;
}
