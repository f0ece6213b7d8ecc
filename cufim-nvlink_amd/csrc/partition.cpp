// partition.cpp -- element partitions (reference tags or own recursive
// bisection) and the device pack built from them.
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>

#include "host.hpp"

namespace fim {

namespace {

// local copy numbering: order of first appearance over the submesh's
// elements (in submesh element order, sorted node order).  For tag
// partitions this is the reference numbering (src/core/types.cpp:391-408).
void number_copies(const Mesh &m, Partition &p) {
  std::vector<int64_t> stamp(m.nn, -1);
  p.s_vdsp.assign(p.nsub + 1, 0);
  p.l2g.clear();
  p.l2g.reserve(m.ne);
  for (int64_t s = 0; s < p.nsub; ++s) {
    p.s_vdsp[s] = (int64_t)p.l2g.size();
    for (int64_t i = p.s_edsp[s]; i < p.s_edsp[s + 1]; ++i) {
      const int64_t e = p.s_elems[i];
      for (int k = 0; k < 4; ++k) {
        const int64_t v = m.e2n[4 * e + k];
        if (stamp[v] != s) {
          stamp[v] = s;
          p.l2g.push_back(v);
        }
      }
    }
  }
  p.s_vdsp[p.nsub] = (int64_t)p.l2g.size();
}

void group_elements(const Mesh &m, Partition &p) {
  p.s_edsp.assign(p.nsub + 1, 0);
  for (int64_t e = 0; e < m.ne; ++e) p.s_edsp[p.e2s[e] + 1]++;
  for (int64_t s = 0; s < p.nsub; ++s) p.s_edsp[s + 1] += p.s_edsp[s];
  p.s_elems.assign(m.ne, 0);
  std::vector<int64_t> fill(p.nsub, 0);
  for (int64_t e = 0; e < m.ne; ++e) p.s_elems[p.s_edsp[p.e2s[e]] + fill[p.e2s[e]]++] = e;
}

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace

// Submesh id = tag in order of first appearance over the elements
// (src/core/types.cpp:329-346).
Partition partition_from_tags(const Mesh &m, const std::vector<int32_t> &tags) {
  if ((int64_t)tags.size() != m.ne)
    throw Error(FIM_EINVAL, "partition tags: expected one tag per element");
  Partition p;
  std::map<int32_t, int64_t> id;
  p.e2s.resize(m.ne);
  for (int64_t e = 0; e < m.ne; ++e) {
    auto it = id.find(tags[e]);
    if (it == id.end()) it = id.emplace(tags[e], (int64_t)id.size()).first;
    p.e2s[e] = it->second;
  }
  p.nsub = (int64_t)id.size();
  group_elements(m, p);
  number_copies(m, p);
  return p;
}

// Recursive coordinate bisection of element centroids.  A part is split
// until it holds at most max_nn vertex copies and max_ne elements; the
// split point divides the elements in the ratio of the number of leaves the
// two sides are expected to need, so leaves come out close to the budget.
// Ties on the split coordinate are broken by a seeded hash of the element
// id, then the id: the result is a pure function of (mesh, budgets, seed).
// Leaves are emitted depth first (spatially coherent submesh order).
Partition partition_rcb(const Mesh &m, int max_nn, int max_ne, uint64_t seed) {
  if (max_nn < 8) throw Error(FIM_EINVAL, "max_nn too small");
  if (max_ne <= 0) max_ne = 8 * max_nn;
  std::vector<double> cen(3 * m.ne);
  for (int64_t e = 0; e < m.ne; ++e)
    for (int d = 0; d < 3; ++d) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += m.xyz[3 * m.e2n[4 * e + k] + d];
      cen[3 * e + d] = 0.25 * s;
    }
  std::vector<uint64_t> key(m.ne);
  for (int64_t e = 0; e < m.ne; ++e) key[e] = mix64(seed ^ mix64((uint64_t)e + 1));
  // connected components (elements sharing a vertex), numbered by their
  // lowest element: bisection never mixes components, so a union of meshes
  // partitions as its parts would on their own
  std::vector<int64_t> uf(m.nn);
  std::iota(uf.begin(), uf.end(), 0);
  auto find = [&](int64_t v) {
    while (uf[v] != v) v = uf[v] = uf[uf[v]];
    return v;
  };
  for (int64_t e = 0; e < m.ne; ++e)
    for (int k = 1; k < 4; ++k) {
      const int64_t a = find(m.e2n[4 * e]), b = find(m.e2n[4 * e + k]);
      if (a != b) uf[std::max(a, b)] = std::min(a, b);
    }
  std::vector<int64_t> comp_of_root(m.nn, -1), ecomp(m.ne);
  int64_t ncomp = 0;
  for (int64_t e = 0; e < m.ne; ++e) {
    const int64_t r = find(m.e2n[4 * e]);
    if (comp_of_root[r] < 0) comp_of_root[r] = ncomp++;
    ecomp[e] = comp_of_root[r];
  }
  std::vector<int64_t> order(m.ne);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t x, int64_t y) { return ecomp[x] < ecomp[y]; });
  std::vector<int64_t> stamp(m.nn, -1);
  int64_t stamp_id = 0;
  Partition p;
  p.e2s.assign(m.ne, -1);
  p.nsub = 0;
  // explicit stack of [begin, end) ranges of `order`, one per component to
  // start, first component on top
  std::vector<std::pair<int64_t, int64_t>> stack;
  for (int64_t i = m.ne; i > 0;) {
    int64_t j = i - 1;
    while (j > 0 && ecomp[order[j - 1]] == ecomp[order[i - 1]]) --j;
    stack.push_back({j, i});
    i = j;
  }
  while (!stack.empty()) {
    auto [b, e] = stack.back();
    stack.pop_back();
    const int64_t cnt = e - b;
    int64_t nv = 0;
    ++stamp_id;
    for (int64_t i = b; i < e; ++i)
      for (int k = 0; k < 4; ++k) {
        const int64_t v = m.e2n[4 * order[i] + k];
        if (stamp[v] != stamp_id) {
          stamp[v] = stamp_id;
          ++nv;
        }
      }
    if ((nv <= max_nn && cnt <= max_ne) || cnt <= 1) {
      for (int64_t i = b; i < e; ++i) p.e2s[order[i]] = p.nsub;
      ++p.nsub;
      continue;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = b; i < e; ++i)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], cen[3 * order[i] + d]);
        hi[d] = std::max(hi[d], cen[3 * order[i] + d]);
      }
    int ax = 0;
    for (int d = 1; d < 3; ++d)
      if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
    // leaves needed: by vertex budget (with headroom) and element budget
    const double fill = 0.85;
    int64_t need = std::max<int64_t>(
        (int64_t)std::ceil((double)nv / (fill * max_nn)),
        (int64_t)std::ceil((double)cnt / (fill * max_ne)));
    need = std::max<int64_t>(need, 2);
    const int64_t left = need / 2;
    int64_t mid = b + (int64_t)((double)cnt * (double)left / (double)need);
    mid = std::min(std::max(mid, b + 1), e - 1);
    auto less = [&](int64_t x, int64_t y) {
      const double cx = cen[3 * x + ax], cy = cen[3 * y + ax];
      if (cx != cy) return cx < cy;
      if (key[x] != key[y]) return key[x] < key[y];
      return x < y;
    };
    std::nth_element(order.begin() + b, order.begin() + mid, order.begin() + e, less);
    // keep each side in ascending element order (locality of the input
    // numbering inside a leaf)
    std::sort(order.begin() + b, order.begin() + mid);
    std::sort(order.begin() + mid, order.begin() + e);
    stack.push_back({mid, e});
    stack.push_back({b, mid});
  }
  group_elements(m, p);
  number_copies(m, p);
  return p;
}

int64_t converge_lds_bytes(int32_t nn_max, int32_t ne_max) {
  // same carve as the default converge-mode kernel: vertex times and
  // candidate minima (2 x f32 per vertex, rounded to 64), four vertex
  // bitmasks, two element bitmasks, element list (u16), source list (u16),
  // 8 counters and 4 region factors, interface-copy bits and entry counts
  auto al = [](int64_t x) { return (x + 15) & ~int64_t(15); };
  const int64_t nn64 = (nn_max + 63) & ~int64_t(63);
  const int64_t ne32 = (ne_max + 31) / 32;
  // (the source list doubles as the (element, target) pair rings of the
  // four waves of a workgroup of 256: at least 4 x 320 u16)
  return 2 * al(4 * nn64) + 4 * al(4 * (nn64 / 32)) + 2 * al(4 * ne32) +
         al(2LL * std::min<int64_t>(ne_max, kElistCap)) +
         al(std::max<int64_t>(2 * nn64, ne_max < (1 << 14) ? 2560 : 0)) + 48 + 2 * al(4 * (nn64 / 32));
}

Pack build_pack(const Mesh &m, const Partition &p, int32_t source) {
  Pack k;
  k.source = source;
  k.part = p;
  k.nn_global = m.nn;
  k.ne_global = m.ne;
  k.nsub = p.nsub;
  k.ncopy = (int64_t)p.l2g.size();
  if (k.ncopy >= (int64_t)UINT32_MAX || m.ne >= (int64_t)UINT32_MAX)
    throw Error(FIM_ELIMIT, "mesh too large for 32-bit copy ids");
  k.front_width = front_width(m);
  k.sub.resize(p.nsub);
  k.e2n_l.resize(4 * m.ne);
  k.geo32.resize(6 * m.ne);
  k.geo64.resize(6 * m.ne);
  k.hgt32.resize(4 * m.ne);
  k.ereg.assign(m.ne, 0);
  k.c2g.resize(k.ncopy);
  k.c2s.resize(k.ncopy);
  std::vector<int64_t> g2l(m.nn, -1);
  for (int64_t s = 0; s < p.nsub; ++s) {
    SubInfo &si = k.sub[s];
    const int64_t nn_s = p.s_vdsp[s + 1] - p.s_vdsp[s];
    const int64_t ne_s = p.s_edsp[s + 1] - p.s_edsp[s];
    if (nn_s > 65535 || ne_s > 65535)
      throw Error(FIM_ELIMIT, "submesh exceeds 65535 vertices or elements");
    si.eoff = (uint32_t)p.s_edsp[s];
    si.ne = (uint32_t)ne_s;
    si.coff = (uint32_t)p.s_vdsp[s];
    si.nn = (uint32_t)nn_s;
    k.nn_max = std::max<int32_t>(k.nn_max, (int32_t)nn_s);
    k.ne_max = std::max<int32_t>(k.ne_max, (int32_t)ne_s);
    for (int64_t l = 0; l < nn_s; ++l) {
      const int64_t c = p.s_vdsp[s] + l;
      g2l[p.l2g[c]] = l;
      k.c2g[c] = (uint32_t)p.l2g[c];
      k.c2s[c] = (uint32_t)s;
    }
    for (int64_t i = p.s_edsp[s]; i < p.s_edsp[s + 1]; ++i) {
      const int64_t e = p.s_elems[i];
      // local ids of the element, sorted ascending (the reference sorts the
      // submesh's elements by local id, src/core/types.cpp:473)
      int64_t loc[4];
      for (int q = 0; q < 4; ++q) loc[q] = g2l[m.e2n[4 * e + q]];
      std::sort(loc, loc + 4);
      for (int q = 0; q < 4; ++q) k.e2n_l[4 * i + q] = (uint16_t)loc[q];
      // geometry in that node order
      double M[6], geo[6];
      metric_tensor(&m.fib[3 * e], m.vel[2 * e], m.vel[2 * e + 1], M);
      static const int pa[6] = {0, 0, 1, 0, 1, 2};
      static const int pb[6] = {1, 2, 2, 3, 3, 3};
      for (int q = 0; q < 6; ++q) {
        const double *A = &m.xyz[3 * p.l2g[p.s_vdsp[s] + loc[pa[q]]]];
        const double *B = &m.xyz[3 * p.l2g[p.s_vdsp[s] + loc[pb[q]]]];
        const double x = A[0] - B[0], y = A[1] - B[1], z = A[2] - B[2];
        geo[q] = (M[0] * x + M[1] * y + M[2] * z) * x + (M[1] * x + M[3] * y + M[4] * z) * y +
                 (M[2] * x + M[4] * y + M[5] * z) * z;
        k.geo64[6 * i + q] = geo[q];
        k.geo32[6 * i + q] = (float)geo[q];
      }
      // heights: for target t with the others relabelled 1,2,3 (A = x1-x3,
      // B = x2-x3, u = x4-x3), h^2 = |u|^2 - g^T G^-1 g with the Gram matrix G
      // of (A, B) and g = (A.u, B.u).  Every candidate of t is at least
      // min(other times) + h, which lets the kernel skip hopeless targets.
      for (int t = 0; t < 4; ++t) {
        const int q1 = (t + 1) & 3, q2 = (t + 2) & 3, q3 = (t + 3) & 3;
        auto E = [&](int a, int b) { return geo[(a > b) ? ((a == 3) ? 3 + b : a + b - 1)
                                                        : ((b == 3) ? 3 + a : a + b - 1)]; };
        const double e12 = E(q1, q2), e13 = E(q1, q3), e23 = E(q2, q3);
        const double e14 = E(q1, t), e24 = E(q2, t), e34 = E(q3, t);
        const double AA = e13, BB = e23, AB = 0.5 * (e13 + e23 - e12);
        const double Au = 0.5 * (e13 + e34 - e14), Bu = 0.5 * (e23 + e34 - e24);
        const double det = AA * BB - AB * AB;
        double h2 = 0.0;
        if (det > 0.0) h2 = e34 - (BB * Au * Au - 2.0 * AB * Au * Bu + AA * Bu * Bu) / det;
        // relative safety margin far above the rounding of h2 for any tet
        // that is not a sliver; slivers get h = 0 (no pruning)
        const double hh = h2 > 1e-6 * e34 ? std::sqrt(h2) * (1.0 - 1e-4) : 0.0;
        k.hgt32[4 * i + t] = (float)hh;
      }
      if (!m.region.empty()) k.ereg[i] = m.region[e];
    }
    for (int64_t l = 0; l < nn_s; ++l) g2l[p.l2g[p.s_vdsp[s] + l]] = -1;
  }
  // global vertex -> copies, ascending submesh
  k.g2c_dsp.assign(m.nn + 1, 0);
  for (int64_t c = 0; c < k.ncopy; ++c) k.g2c_dsp[k.c2g[c] + 1]++;
  for (int64_t v = 0; v < m.nn; ++v) k.g2c_dsp[v + 1] += k.g2c_dsp[v];
  k.g2c.assign(k.ncopy, 0);
  {
    std::vector<uint32_t> fill(m.nn, 0);
    for (int64_t c = 0; c < k.ncopy; ++c) {
      const uint32_t v = k.c2g[c];
      k.g2c[k.g2c_dsp[v] + fill[v]++] = (uint32_t)c;
    }
  }
  for (int64_t v = 0; v < m.nn; ++v) {
    const uint32_t n = k.g2c_dsp[v + 1] - k.g2c_dsp[v];
    if (n == 0) throw Error(FIM_EINVAL, "vertex not used by any element");
    if (n >= 2) k.grp_v.push_back((uint32_t)v);
  }
  // interface lists per submesh
  k.if_dsp.push_back(0);
  for (int64_t s = 0; s < p.nsub; ++s) {
    SubInfo &si = k.sub[s];
    si.ioff = (uint32_t)k.if_l.size();
    for (uint32_t l = 0; l < si.nn; ++l) {
      const uint32_t c = si.coff + l, v = k.c2g[c];
      const uint32_t b = k.g2c_dsp[v], e = k.g2c_dsp[v + 1];
      if (e - b < 2) continue;
      k.if_l.push_back((uint16_t)l);
      for (uint32_t j = b; j < e; ++j)
        if (k.g2c[j] != c) k.if_other.push_back(k.g2c[j]);
      k.if_dsp.push_back((uint32_t)k.if_other.size());
    }
    si.ni = (uint32_t)k.if_l.size() - si.ioff;
  }
  // compat CSR: per copy incident local elements and neighbour copies
  k.n2e_dsp.assign(k.ncopy + 1, 0);
  k.n2n_dsp.assign(k.ncopy + 1, 0);
  for (int64_t s = 0; s < p.nsub; ++s) {
    const SubInfo &si = k.sub[s];
    std::vector<std::vector<uint16_t>> nb(si.nn);
    // count per copy
    std::vector<uint32_t> cnt(si.nn, 0);
    for (uint32_t le = 0; le < si.ne; ++le)
      for (int q = 0; q < 4; ++q) cnt[k.e2n_l[4 * (si.eoff + le) + q]]++;
    const uint32_t base = (uint32_t)k.n2e_le.size();
    std::vector<uint32_t> dsp(si.nn + 1, 0);
    for (uint32_t l = 0; l < si.nn; ++l) dsp[l + 1] = dsp[l] + cnt[l];
    k.n2e_le.resize(base + dsp[si.nn]);
    k.n2e_pos.resize(base + dsp[si.nn]);
    std::vector<uint32_t> fill(si.nn, 0);
    for (uint32_t le = 0; le < si.ne; ++le)
      for (int q = 0; q < 4; ++q) {
        const uint16_t l = k.e2n_l[4 * (si.eoff + le) + q];
        const uint32_t at = base + dsp[l] + fill[l]++;
        k.n2e_le[at] = (uint16_t)le;
        k.n2e_pos[at] = (uint8_t)q;
        for (int r = 0; r < 4; ++r) nb[l].push_back(k.e2n_l[4 * (si.eoff + le) + r]);
      }
    for (uint32_t l = 0; l < si.nn; ++l) {
      k.n2e_dsp[si.coff + l] = base + dsp[l];
      k.n2e_dsp[si.coff + l + 1] = base + dsp[l + 1];
      std::sort(nb[l].begin(), nb[l].end());
      nb[l].erase(std::unique(nb[l].begin(), nb[l].end()), nb[l].end());
      k.n2n_dsp[si.coff + l] = (uint32_t)k.n2n_lv.size();
      k.n2n_lv.insert(k.n2n_lv.end(), nb[l].begin(), nb[l].end());
      k.n2n_dsp[si.coff + l + 1] = (uint32_t)k.n2n_lv.size();
    }
  }
  return k;
}

}  // namespace fim
