// local_solve.hpp -- per-element local eikonal updates, host and device.
//
// Element geometry is the 6 squared metric edge lengths e_ij^T M e_ij in the
// element's node order (e12 e13 e23 e14 e24 e34).  A candidate for node
// position T uses the other three nodes relabelled cyclically as 1, 2, 3 and
// the target as 4 -- the relabelling the reference encodes in its tables
// (src/eikonal/cuda_fim.h:41-48, applied at 1384-1392).
#pragma once

#include <cfloat>
#include <cmath>

#ifndef FIM_HD
#ifdef __HIPCC__
#define FIM_HD __host__ __device__
#else
#define FIM_HD
#endif
#endif

namespace fim {

// square root and quotient of the exact candidates.  FP64 (and host code)
// uses the correctly rounded operations; the FP32 device path uses the
// hardware square root and reciprocal (1 ulp), which is far below the FP32
// solver's error budget.
FIM_HD inline double csqrt(double x) { return std::sqrt(x); }
FIM_HD inline double cdiv(double a, double b) { return a / b; }
FIM_HD inline float csqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return std::sqrt(x);
#endif
}
FIM_HD inline float cdiv(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return a * __builtin_amdgcn_rcpf(b);
#else
  return a / b;
#endif
}

// index of the edge between node positions i, j in the 6-entry geometry
FIM_HD constexpr int edge_idx(int i, int j) {
  return (i > j) ? edge_idx(j, i) : (j == 3 ? 3 + i : i + j - 1);
}

// ---------------------------------------------------------------------------
// exact candidates (converge mode, FP32).  Minimum over the edge / face of
// interpolated arrival time plus metric distance to the target.
//   edge (1,2), target 4:  p = x1 + l (x2 - x1),  G = |e12|^2,
//     g = e12.e14 = (|e12|^2 + |e14|^2 - |e24|^2)/2,  d = p2 - p1,
//     s^2 (1 - d^2/G) = |e14|^2 - g^2/G,  l = (g - s d)/G  in (0,1).
//   face (1,2,3), target 4: p = x3 + l1 (x1-x3) + l2 (x2-x3), Gram matrix G
//     of (x1-x3, x2-x3), g = their products with x4-x3, d = (p1-p3, p2-p3),
//     s^2 (1 - d^T G^-1 d) = |e34|^2 - g^T G^-1 g,  l = G^-1 (g - s d).
// ---------------------------------------------------------------------------
// Both solvers are evaluated without branches: every quantity is computed
// and an invalid configuration (degenerate edge / face, ray leaving the
// simplex, unusable node time) selects `inf` at the end.  Where the
// configuration is valid the operations are exactly the sequential ones.
template <class T>
FIM_HD inline T tri_exact(T e12, T e14, T e24, T p1, T p2, T inf) {
  const T G = e12, g = T(0.5) * (e12 + e14 - e24), d = p2 - p1;
  const T den = G - d * d;
  const T r = e14 - cdiv(g * g, G);
  const T s = csqrt(std::fmax(cdiv(r * G, den), T(0)));
  const T l = cdiv(g - s * d, G);
  const bool ok = (den > T(0)) && (r > T(0)) && (l > T(0)) && (l < T(1));
  return ok ? p1 + l * d + s : inf;
}

template <class T>
FIM_HD inline T tet_exact(T e12, T e13, T e23, T e14, T e24, T e34, T p1, T p2, T p3,
                          T inf) {
  const T AA = e13, BB = e23, AB = T(0.5) * (e13 + e23 - e12);
  const T Au = T(0.5) * (e13 + e34 - e14), Bu = T(0.5) * (e23 + e34 - e24);
  const T d1 = p1 - p3, d2 = p2 - p3;
  const T det = AA * BB - AB * AB;
  const T w = cdiv(BB * d1 * d1 - T(2) * AB * d1 * d2 + AA * d2 * d2, det);
  const T r = e34 - cdiv(BB * Au * Au - T(2) * AB * Au * Bu + AA * Bu * Bu, det);
  const T s = csqrt(std::fmax(cdiv(r, T(1) - w), T(0)));
  const T t1 = Au - s * d1, t2 = Bu - s * d2;
  const T l1 = cdiv(BB * t1 - AB * t2, det);
  const T l2 = cdiv(AA * t2 - AB * t1, det);
  const bool ok = (det > T(0)) && (w < T(1)) && (r > T(0)) && (l1 > T(0)) && (l2 > T(0)) &&
                  (l1 + l2 < T(1));
  return ok ? p3 + l1 * d1 + l2 * d2 + s : inf;
}

// candidate of node 4 from nodes 1, 2, 3 (already relabelled); a time >= inf
// marks an unusable node
template <class T>
FIM_HD inline T cand_core(T e12, T e13, T e23, T e14, T e24, T e34, T p1, T p2, T p3, T inf) {
  const bool u1 = p1 < inf, u2 = p2 < inf, u3 = p3 < inf;
  T best = std::fmin(std::fmin(u1 ? p1 + csqrt(e14) : inf, u2 ? p2 + csqrt(e24) : inf),
                     u3 ? p3 + csqrt(e34) : inf);
  const T t12 = tri_exact<T>(e12, e14, e24, p1, p2, inf);
  const T t13 = tri_exact<T>(e13, e14, e34, p1, p3, inf);
  const T t23 = tri_exact<T>(e23, e24, e34, p2, p3, inf);
  const T t123 = tet_exact<T>(e12, e13, e23, e14, e24, e34, p1, p2, p3, inf);
  best = std::fmin(best, (u1 && u2) ? t12 : inf);
  best = std::fmin(best, (u1 && u3) ? t13 : inf);
  best = std::fmin(best, (u2 && u3) ? t23 : inf);
  best = std::fmin(best, (u1 && u2 && u3) ? t123 : inf);
  return best;
}

// tri_exact with the reciprocal of G = e12 given
template <class T>
FIM_HD inline T tri_exact_r(T e12, T rG, T e14, T e24, T p1, T p2, T inf) {
  const T G = e12, g = T(0.5) * (e12 + e14 - e24), d = p2 - p1;
  const T den = G - d * d;
  const T r = e14 - g * g * rG;
  const T s = csqrt(std::fmax(cdiv(r * G, den), T(0)));
  const T l = (g - s * d) * rG;
  const bool ok = (den > T(0)) && (r > T(0)) && (l > T(0)) && (l < T(1));
  return ok ? p1 + l * d + s : inf;
}

// candidate for node position TP with per-element square roots sq[k] and
// reciprocals rc[k] of the six squared lengths
template <int TP, class T>
FIM_HD inline T cand_exact_sh(const T g[6], const T sq[6], const T rc[6], const T ph[4], T inf) {
  constexpr int q1 = (TP + 1) & 3, q2 = (TP + 2) & 3, q3 = (TP + 3) & 3;
  constexpr int i12 = edge_idx(q1, q2), i13 = edge_idx(q1, q3), i23 = edge_idx(q2, q3);
  constexpr int i14 = edge_idx(q1, TP), i24 = edge_idx(q2, TP), i34 = edge_idx(q3, TP);
  const T p1 = ph[q1], p2 = ph[q2], p3 = ph[q3];
  const bool u1 = p1 < inf, u2 = p2 < inf, u3 = p3 < inf;
  T best = std::fmin(std::fmin(u1 ? p1 + sq[i14] : inf, u2 ? p2 + sq[i24] : inf),
                     u3 ? p3 + sq[i34] : inf);
  const T t12 = tri_exact_r<T>(g[i12], rc[i12], g[i14], g[i24], p1, p2, inf);
  const T t13 = tri_exact_r<T>(g[i13], rc[i13], g[i14], g[i34], p1, p3, inf);
  const T t23 = tri_exact_r<T>(g[i23], rc[i23], g[i24], g[i34], p2, p3, inf);
  const T t123 = tet_exact<T>(g[i12], g[i13], g[i23], g[i14], g[i24], g[i34], p1, p2, p3, inf);
  best = std::fmin(best, (u1 && u2) ? t12 : inf);
  best = std::fmin(best, (u1 && u3) ? t13 : inf);
  best = std::fmin(best, (u2 && u3) ? t23 : inf);
  best = std::fmin(best, (u1 && u2 && u3) ? t123 : inf);
  return best;
}

// candidate for node position TP (compile time)
template <int TP, class T>
FIM_HD inline T cand_exact(const T g[6], const T ph[4], T inf) {
  constexpr int q1 = (TP + 1) & 3, q2 = (TP + 2) & 3, q3 = (TP + 3) & 3;
  return cand_core<T>(g[edge_idx(q1, q2)], g[edge_idx(q1, q3)], g[edge_idx(q2, q3)],
                      g[edge_idx(q1, TP)], g[edge_idx(q2, TP)], g[edge_idx(q3, TP)], ph[q1], ph[q2],
                      ph[q3], inf);
}

// ---------------------------------------------------------------------------
// reference update (compat mode, FP64 with the reference's float
// intermediates), src/eikonal/cuda_fim.h:1256-1411.  Unusable nodes carry
// DBL_MAX, the reference's "infinity".  Evaluated without FMA contraction so
// results are bit-identical to the CPU restatement.
// ---------------------------------------------------------------------------
FIM_HD inline double tri_ref(double e12, double e13, double e23, double phi1, double phi2) {
#pragma clang fp contract(off)
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
  return DBL_MAX;
}

// The reference tet candidate evaluates a closed form whose coefficients are
// named inconsistently (src/eikonal/cuda_fim.h:1307-1312 vs 1352-1357); the
// expression is kept as the reference evaluates it so compat results match.
FIM_HD inline double tet_ref(double e13, double e23, double e34, double e12, double e14,
                             double e24, double phi1, double phi2, double phi3) {
#pragma clang fp contract(off)
  const double a = e13;
  const double b = 0.5 * (e23 + e13 - e12);
  const double c = -0.5 * (e34 + e13 - e14);
  const double d = e23;
  const double e = -0.5 * (e34 + e23 - e24);
  const double f = e34;
  const double p13 = phi3 - phi1;
  const double p23 = phi3 - phi2;
  const double sqt =
      -e * e * c * d * d + a * d * d * d * d - (b * b * c - a * c * c) * f * f +
      ((b * b - a * c) * f * f - (e * e * c - a * d * d - 2. * (e * c - b * d) * e) * f) * p13 * p13 -
      2. * (e * e * c * d - 2. * e * b * d * d + a * d * d * d + (b * b - a * c) * d * f) * p13 * p23 +
      (e * e * c * c - 2. * e * b * c * d + a * c * d * d + (b * b * c - a * c * c) * f) * p23 * p23 +
      2. * (e * c * d * d - b * d * d * d) * e +
      (e * e * c * c + (b * b - 2. * a * c) * d * d - 2. * (e * c * c - b * c * d) * e) * f;
  if (sqt < 0.) return DBL_MAX;
  const double q1 = e * c * d * d - b * d * d * d + (e * c - b * d) * f * p13 * p13 -
                    2. * (e * c * d - b * d * d) * p13 * p23 +
                    (e * c * c - b * c * d) * p23 * p23 - (e * c * c - b * c * d) * f;
  const double q2 = (d * d * d * d - 2. * c * d * d * f + c * c * f * f +
                     (d * d * f - c * f * f) * p13 * p13 -
                     2. * (d * d * d - c * d * f) * p13 * p23 +
                     (c * d * d - c * c * f) * p23 * p23);
  const double m1 = (q1 + sqt * (d * p13 - c * p23)) / q2;
  const double m2 = (q1 - sqt * (d * p13 - c * p23)) / q2;
  const double k1 = -((f * m1 + e) * p13 - (d * m1 + b) * p23) / (d * p13 - c * p23);
  const double k2 = -((f * m2 + e) * p13 - (d * m2 + b) * p23) / (d * p13 - c * p23);
  double w = 1 - k1 - m1, r1 = DBL_MAX, r2 = DBL_MAX;
  if ((k1 > 0. && k1 < 1.) && (m1 > 0. && m1 < 1.) && w > 0.)
    r1 = k1 * phi1 + m1 * phi2 + w * phi3 +
         sqrt(a + 2. * k1 * b + k1 * k1 * c + 2. * k1 * m1 * d + 2. * m1 * e + m1 * m1 * f);
  w = 1 - k2 - m2;
  if ((k2 > 0. && k2 < 1.) && (m2 > 0. && m2 < 1.) && w > 0.)
    r2 = k2 * phi1 + m2 * phi2 + w * phi3 +
         sqrt(a + 2. * k2 * b + k2 * k2 * c + 2. * k2 * m2 * d + 2. * m2 * e + m2 * m2 * f);
  return fmin(r1, r2);
}

// reference candidate for node position t (runtime), after the upwind filter
FIM_HD inline double cand_ref(const double g[6], const double ph[4], int t) {
#pragma clang fp contract(off)
  const int q1 = (t + 1) & 3, q2 = (t + 2) & 3, q3 = (t + 3) & 3;
  const double e13 = g[edge_idx(q1, q3)], e23 = g[edge_idx(q2, q3)], e34 = g[edge_idx(q3, t)];
  const double e12 = g[edge_idx(q1, q2)], e14 = g[edge_idx(q1, t)], e24 = g[edge_idx(q2, t)];
  const double p1 = ph[q1], p2 = ph[q2], p3 = ph[q3];
  double v = tet_ref(e13, e23, e34, e12, e14, e24, p1, p2, p3);
  v = fmin(v, tri_ref(e12, e14, e24, p1, p2));
  v = fmin(v, tri_ref(e23, e24, e34, p2, p3));
  v = fmin(v, tri_ref(e13, e14, e34, p1, p3));
  v = fmin(v, p1 + sqrt(e14));
  v = fmin(v, p2 + sqrt(e24));
  v = fmin(v, p3 + sqrt(e34));
  return v;
}

}  // namespace fim
