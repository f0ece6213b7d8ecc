// mesh.cpp -- mesh container, connectivity, metric geometry, legacy-VTK I/O,
// partition tags and the synthetic box generator.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "host.hpp"

namespace fim {

void Mesh::finalize() {
  if (nn <= 0 || ne <= 0) throw Error(FIM_EINVAL, "mesh has no vertices or no elements");
  if ((int64_t)xyz.size() != 3 * nn || (int64_t)e2n.size() != 4 * ne)
    throw Error(FIM_EINVAL, "mesh array sizes inconsistent");
  if (fib.empty()) {
    fib.assign(3 * ne, 0.0);
    for (int64_t e = 0; e < ne; ++e) fib[3 * e] = 1.0;
  }
  if (vel.empty()) vel.assign(2 * ne, 0.6);
  if (etags.empty()) etags.assign(ne, 1);
  for (int64_t e = 0; e < ne; ++e) {
    int64_t *c = e2n.data() + 4 * e;
    std::sort(c, c + 4);
    for (int k = 0; k < 4; ++k)
      if (c[k] < 0 || c[k] >= nn) throw Error(FIM_EINVAL, "element node id out of range");
    for (int k = 1; k < 4; ++k)
      if (c[k] == c[k - 1]) throw Error(FIM_EINVAL, "degenerate element (repeated node)");
  }
  for (int64_t e = 0; e < ne; ++e)
    if (!(vel[2 * e] > 0.0) || !(vel[2 * e + 1] > 0.0))
      throw Error(FIM_EINVAL, "conduction velocities must be positive");
  n2e_dsp.assign(nn + 1, 0);
  for (int64_t i = 0; i < 4 * ne; ++i) n2e_dsp[e2n[i] + 1]++;
  for (int64_t v = 0; v < nn; ++v) n2e_dsp[v + 1] += n2e_dsp[v];
  n2e_con.assign(4 * ne, 0);
  std::vector<int64_t> fill(nn, 0);
  for (int64_t e = 0; e < ne; ++e)
    for (int k = 0; k < 4; ++k) {
      const int64_t v = e2n[4 * e + k];
      n2e_con[n2e_dsp[v] + fill[v]++] = e;
    }
}

// Same algebra and operation order as compute_velocity_tensor_f3
// (src/core/types.h:174-188): host geometry is reproduced bit for bit.
void metric_tensor(const double f[3], double vfs, double vss, double M[6]) {
  const double fa = f[0] * f[0], fb = f[0] * f[1], fc = f[0] * f[2];
  const double fd = f[1] * f[1], fe = f[1] * f[2], ff = f[2] * f[2];
  const double vf = 1. / vfs / vfs, vs = 1. / vss / vss;
  M[0] = vf * fa + vs * (1. - fa);
  M[1] = (vf - vs) * fb;
  M[2] = (vf - vs) * fc;
  M[3] = vf * fd + vs * (1. - fd);
  M[4] = (vf - vs) * fe;
  M[5] = vf * ff + vs * (1. - ff);
}

void element_geo(const Mesh &m, int64_t e, double geo[6]) {
  double M[6];
  metric_tensor(&m.fib[3 * e], m.vel[2 * e], m.vel[2 * e + 1], M);
  static const int pa[6] = {0, 0, 1, 0, 1, 2};
  static const int pb[6] = {1, 2, 2, 3, 3, 3};
  const int64_t *c = &m.e2n[4 * e];
  for (int k = 0; k < 6; ++k) {
    const double *A = &m.xyz[3 * c[pa[k]]], *B = &m.xyz[3 * c[pb[k]]];
    const double x = A[0] - B[0], y = A[1] - B[1], z = A[2] - B[2];
    geo[k] = (M[0] * x + M[1] * y + M[2] * z) * x + (M[1] * x + M[3] * y + M[4] * z) * y +
             (M[2] * x + M[4] * y + M[5] * z) * z;
  }
}

// 1.01 * longest edge / 0.6 (src/eikonal/cuda_fim.h:1536-1576)
double front_width(const Mesh &m) {
  double best = 0.0;
  for (int64_t e = 0; e < m.ne; ++e)
    for (int i = 0; i < 4; ++i)
      for (int j = i + 1; j < 4; ++j) {
        const double *A = &m.xyz[3 * m.e2n[4 * e + i]], *B = &m.xyz[3 * m.e2n[4 * e + j]];
        const double x = B[0] - A[0], y = B[1] - A[1], z = B[2] - A[2];
        best = std::max(best, x * x + y * y + z * z);
      }
  return 1.01 * std::sqrt(best) / 0.6;
}

// ---------------------------------------------------------------------------
// legacy VTK, ASCII, tetrahedra (the subset the reference reads:
// src/io/vtk.cpp:11-105)
// ---------------------------------------------------------------------------
Mesh load_vtk(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw Error(FIM_EIO, "file not found: " + path);
  Mesh m;
  std::string line;
  for (int i = 0; i < 4; ++i)
    if (!std::getline(f, line)) throw Error(FIM_EIO, "truncated VTK header: " + path);
  bool have_pts = false, have_cells = false;
  while (std::getline(f, line)) {
    std::istringstream ls(line);
    std::string key;
    if (!(ls >> key)) continue;
    if (key == "POINTS") {
      int64_t n = 0;
      ls >> n;
      if (n <= 0) throw Error(FIM_EIO, "bad POINTS count");
      m.nn = n;
      m.xyz.resize(3 * n);
      for (int64_t i = 0; i < 3 * n; ++i)
        if (!(f >> m.xyz[i])) throw Error(FIM_EIO, "truncated POINTS");
      have_pts = true;
    } else if (key == "CELLS") {
      int64_t n = 0;
      ls >> n;
      if (n <= 0) throw Error(FIM_EIO, "bad CELLS count");
      m.ne = n;
      m.e2n.resize(4 * n);
      for (int64_t e = 0; e < n; ++e) {
        int64_t cnt = 0;
        if (!(f >> cnt)) throw Error(FIM_EIO, "truncated CELLS");
        if (cnt != 4) throw Error(FIM_EIO, "only tetrahedra are supported");
        for (int k = 0; k < 4; ++k)
          if (!(f >> m.e2n[4 * e + k])) throw Error(FIM_EIO, "truncated CELLS");
      }
      have_cells = true;
    } else if (key == "CELL_TYPES") {
      int64_t n = 0;
      ls >> n;
      for (int64_t e = 0; e < n; ++e) {
        int t = 0;
        if (!(f >> t)) throw Error(FIM_EIO, "truncated CELL_TYPES");
        if (t != 10) throw Error(FIM_EIO, "only tetrahedra (type 10) are supported");
      }
    } else if (key == "CELL_DATA") {
      int64_t n = 0;
      ls >> n;
      std::string w;
      // SCALARS <name> <type> [1] / LOOKUP_TABLE default
      std::getline(f, line);
      std::getline(f, line);
      m.etags.resize(n);
      for (int64_t e = 0; e < n; ++e) {
        double v;
        if (!(f >> v)) throw Error(FIM_EIO, "truncated CELL_DATA");
        m.etags[e] = (int32_t)v;
      }
    }
  }
  if (!have_pts || !have_cells) throw Error(FIM_EIO, "VTK file lacks POINTS or CELLS: " + path);
  if (!m.etags.empty() && (int64_t)m.etags.size() != m.ne) m.etags.clear();
  m.finalize();
  return m;
}

// byte layout of save_vtk (src/io/vtk.cpp:109-212)
void save_vtk(const Mesh &m, const std::string &path, const double *phi, const int32_t *state) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) throw Error(FIM_EIO, "cannot open file: " + path);
  int64_t nn = 0;
  for (int64_t v : m.e2n) nn = std::max(nn, v + 1);
  std::fprintf(f, "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET UNSTRUCTURED_GRID\n\n");
  std::fprintf(f, "POINTS %d     float\n", (int)nn);
  for (int64_t i = 0; i < nn; ++i) {
    const float x = (float)m.xyz[3 * i], y = (float)m.xyz[3 * i + 1], z = (float)m.xyz[3 * i + 2];
    std::fprintf(f, "%.3f %.3f %.3f\n", x, y, z);
  }
  std::fprintf(f, "CELL_TYPES %d\n", (int)m.ne);
  for (int64_t e = 0; e < m.ne; ++e) std::fprintf(f, "%d\n", 10);
  std::fprintf(f, "CELLS %d %lu\n", (int)m.ne, (unsigned long)(5 * m.ne));
  for (int64_t e = 0; e < m.ne; ++e) {
    const int64_t *c = &m.e2n[4 * e];
    std::fprintf(f, "%d %ld %ld %ld %ld\n", 4, (long)c[0], (long)c[1], (long)c[2], (long)c[3]);
  }
  std::fprintf(f, "CELL_DATA %d \n", (int)m.ne);
  std::fprintf(f, "SCALARS %s int %d\n", "elemTag", 1);
  std::fprintf(f, "LOOKUP_TABLE default\n");
  for (int64_t e = 0; e < m.ne; ++e) std::fprintf(f, "%d \n", m.etags[e]);
  if (phi || state) {
    std::fprintf(f, "POINT_DATA %d \n", (int)nn);
    if (phi) {
      std::fprintf(f, "SCALARS %s double %d\n", "phi", 1);
      std::fprintf(f, "LOOKUP_TABLE default\n");
      for (int64_t i = 0; i < nn; ++i) std::fprintf(f, "%lf \n", (phi[i] < 1e50) ? phi[i] : -1.0);
    }
    if (state) {
      std::fprintf(f, "SCALARS %s double %d\n", "states", 1);
      std::fprintf(f, "LOOKUP_TABLE default\n");
      for (int64_t i = 0; i < nn; ++i) std::fprintf(f, "%lf \n", (double)state[i]);
    }
  }
  std::fclose(f);
}

std::vector<int32_t> read_tags(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw Error(FIM_EIO, "file not found: " + path);
  std::vector<int32_t> t;
  int64_t v;
  while (f >> v) t.push_back((int32_t)v);
  if (!f.eof()) throw Error(FIM_EIO, "non-integer token in tags file: " + path);
  return t;
}

// ---------------------------------------------------------------------------
// synthetic box meshes
// ---------------------------------------------------------------------------
namespace {
struct Xoshiro256ss {
  uint64_t s[4];
  explicit Xoshiro256ss(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {  // SplitMix64 expansion
      uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};
}  // namespace

// Box [0,lx]x[0,ly]x[0,lz], nx*ny*nz hexes, each split into the 6 tets of
// the Kuhn subdivision along its (0,0,0)-(1,1,1) diagonal; the split is the
// same in every hex, so faces match.
Mesh gen_box(double lx, double ly, double lz, int nx, int ny, int nz, double jitter,
             uint64_t seed, double a0, double a1, double vf, double vc) {
  if (nx < 1 || ny < 1 || nz < 1 || !(lx > 0) || !(ly > 0) || !(lz > 0))
    throw Error(FIM_EINVAL, "bad box dimensions");
  if (jitter < 0.0 || jitter >= 0.5) throw Error(FIM_EINVAL, "jitter must be in [0, 0.5)");
  Mesh m;
  const int64_t NX = nx + 1, NY = ny + 1, NZ = nz + 1;
  m.nn = NX * NY * NZ;
  m.ne = 6LL * nx * ny * nz;
  m.xyz.resize(3 * m.nn);
  const double hx = lx / nx, hy = ly / ny, hz = lz / nz;
  Xoshiro256ss rng(seed);
  for (int64_t k = 0; k < NZ; ++k)
    for (int64_t j = 0; j < NY; ++j)
      for (int64_t i = 0; i < NX; ++i) {
        const int64_t v = i + NX * (j + NY * k);
        double x = i * hx, y = j * hy, z = k * hz;
        if (i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz && jitter > 0.0) {
          x += (2.0 * rng.uniform() - 1.0) * jitter * hx;
          y += (2.0 * rng.uniform() - 1.0) * jitter * hy;
          z += (2.0 * rng.uniform() - 1.0) * jitter * hz;
        }
        m.xyz[3 * v] = x;
        m.xyz[3 * v + 1] = y;
        m.xyz[3 * v + 2] = z;
      }
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  m.e2n.resize(4 * m.ne);
  m.fib.resize(3 * m.ne);
  m.vel.resize(2 * m.ne);
  int64_t e = 0;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i)
        for (int p = 0; p < 6; ++p) {
          int64_t c[3] = {i, j, k};
          int64_t *out = &m.e2n[4 * e];
          out[0] = c[0] + NX * (c[1] + NY * c[2]);
          for (int s = 0; s < 3; ++s) {
            c[perm[p][s]]++;
            out[s + 1] = c[0] + NX * (c[1] + NY * c[2]);
          }
          // fibre angle from the element's mean height
          double zc = 0.0;
          for (int s = 0; s < 4; ++s) zc += m.xyz[3 * out[s] + 2];
          zc *= 0.25;
          const double th = (a0 + (a1 - a0) * (zc / lz)) * (M_PI / 180.0);
          m.fib[3 * e] = std::cos(th);
          m.fib[3 * e + 1] = std::sin(th);
          m.fib[3 * e + 2] = 0.0;
          m.vel[2 * e] = vf;
          m.vel[2 * e + 1] = vc;
          ++e;
        }
  m.etags.assign(m.ne, 1);
  m.finalize();
  return m;
}

}  // namespace fim
