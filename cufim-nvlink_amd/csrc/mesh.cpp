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

// Prolate-ellipsoid shell: hexes of side h whose centre lies between the
// inner and outer ellipsoid (semi-axes ai/bi/bi and ao/bo/bo along x, y, z,
// centred at the origin), each split into the same 6 tets as gen_box
// (conforming).  Fibres rotate from +60 deg (inner surface) to -60 deg
// (outer surface) in the tangent plane, from the circumferential direction
// about the long axis towards the longitudinal one; speeds vf / vc, and
// vs inside each of n_scars seeded spheres (radius r0..r1, centres at
// random kept hex centres).  Region id 1 marks scar elements.
Mesh gen_shell(double ao, double bo, double ai, double bi, double h, int n_scars, double r0,
               double r1, uint64_t seed, double vf, double vc, double vs) {
  if (!(h > 0) || !(ai > 0) || !(bi > 0) || !(ao > ai) || !(bo > bi))
    throw Error(FIM_EINVAL, "bad shell dimensions");
  if (n_scars < 0 || !(r1 >= r0) || !(r0 >= 0)) throw Error(FIM_EINVAL, "bad scar parameters");
  const int64_t nx = (int64_t)std::ceil(2 * ao / h), ny = (int64_t)std::ceil(2 * bo / h);
  const int64_t nz = ny;
  if (nx * ny * nz > (int64_t)4e9) throw Error(FIM_ELIMIT, "shell grid too large");
  const double x0 = -0.5 * nx * h, y0 = -0.5 * ny * h, z0 = -0.5 * nz * h;
  auto rr = [](double x, double y, double z, double a, double b) {
    return std::sqrt((x / a) * (x / a) + (y / b) * (y / b) + (z / b) * (z / b));
  };
  auto inside = [&](double x, double y, double z) {
    return rr(x, y, z, ao, bo) <= 1.0 && rr(x, y, z, ai, bi) >= 1.0;
  };
  // kept cells and used grid vertices
  const int64_t NX = nx + 1, NY = ny + 1;
  std::vector<int64_t> cells;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i)
        if (inside(x0 + (i + 0.5) * h, y0 + (j + 0.5) * h, z0 + (k + 0.5) * h))
          cells.push_back(i + nx * (j + ny * k));
  if (cells.empty()) throw Error(FIM_EINVAL, "shell has no cells at this spacing");
  std::vector<int64_t> vid;  // grid vertex -> mesh vertex (sorted map via sort/unique)
  std::vector<int64_t> used;
  used.reserve(cells.size() * 2);
  for (int64_t c : cells) {
    const int64_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) used.push_back((i + di) + NX * ((j + dj) + NY * (k + dk)));
  }
  std::sort(used.begin(), used.end());
  used.erase(std::unique(used.begin(), used.end()), used.end());
  auto vindex = [&](int64_t g) { return (int64_t)(std::lower_bound(used.begin(), used.end(), g) - used.begin()); };
  Mesh m;
  m.nn = (int64_t)used.size();
  m.ne = 6 * (int64_t)cells.size();
  m.xyz.resize(3 * m.nn);
  for (int64_t v = 0; v < m.nn; ++v) {
    const int64_t g = used[v];
    m.xyz[3 * v] = x0 + (g % NX) * h;
    m.xyz[3 * v + 1] = y0 + ((g / NX) % NY) * h;
    m.xyz[3 * v + 2] = z0 + (g / (NX * NY)) * h;
  }
  // scars
  Xoshiro256ss rng(seed);
  std::vector<double> sc(4 * (size_t)n_scars);
  for (int s = 0; s < n_scars; ++s) {
    const int64_t c = cells[(size_t)(rng.uniform() * cells.size()) % cells.size()];
    sc[4 * s] = x0 + (c % nx + 0.5) * h;
    sc[4 * s + 1] = y0 + ((c / nx) % ny + 0.5) * h;
    sc[4 * s + 2] = z0 + (c / (nx * ny) + 0.5) * h;
    sc[4 * s + 3] = r0 + (r1 - r0) * rng.uniform();
  }
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  m.e2n.resize(4 * m.ne);
  m.fib.resize(3 * m.ne);
  m.vel.resize(2 * m.ne);
  m.region.assign(m.ne, 0);
  int64_t e = 0;
  for (int64_t cc : cells) {
    const int64_t ci = cc % nx, cj = (cc / nx) % ny, ck = cc / (nx * ny);
    const double x = x0 + (ci + 0.5) * h, y = y0 + (cj + 0.5) * h, z = z0 + (ck + 0.5) * h;
    // transmural depth: 0 on the inner surface, 1 on the outer
    const double rin = rr(x, y, z, ai, bi), rout = rr(x, y, z, ao, bo);
    const double dep = std::min(1.0, std::max(0.0, (rin - 1.0) / std::max(rin - rout, 1e-12)));
    const double th = (60.0 - 120.0 * dep) * (M_PI / 180.0);
    double n[3] = {x / (ao * ao), y / (bo * bo), z / (bo * bo)};
    double nl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (double &q : n) q /= nl;
    double cr[3] = {0.0, -z, y};
    double cl = std::sqrt(y * y + z * z);
    if (cl < 1e-9) {
      cr[1] = 1.0;
      cr[2] = 0.0;
      cl = 1.0;
    }
    for (double &q : cr) q /= cl;
    // longitudinal tangent = n x c
    double lo[3] = {n[1] * cr[2] - n[2] * cr[1], n[2] * cr[0] - n[0] * cr[2], n[0] * cr[1] - n[1] * cr[0]};
    const double ll = std::sqrt(lo[0] * lo[0] + lo[1] * lo[1] + lo[2] * lo[2]);
    for (double &q : lo) q /= ll;
    double f[3];
    for (int d = 0; d < 3; ++d) f[d] = std::cos(th) * cr[d] + std::sin(th) * lo[d];
    bool scar = false;
    for (int s = 0; s < n_scars && !scar; ++s) {
      const double dx = x - sc[4 * s], dy = y - sc[4 * s + 1], dz = z - sc[4 * s + 2];
      scar = dx * dx + dy * dy + dz * dz <= sc[4 * s + 3] * sc[4 * s + 3];
    }
    for (int p = 0; p < 6; ++p) {
      int64_t c[3] = {ci, cj, ck};
      int64_t *out = &m.e2n[4 * e];
      out[0] = vindex(c[0] + NX * (c[1] + NY * c[2]));
      for (int s = 0; s < 3; ++s) {
        c[perm[p][s]]++;
        out[s + 1] = vindex(c[0] + NX * (c[1] + NY * c[2]));
      }
      for (int d = 0; d < 3; ++d) m.fib[3 * e + d] = f[d];
      m.vel[2 * e] = scar ? vs : vf;
      m.vel[2 * e + 1] = scar ? vs : vc;
      m.region[e] = scar ? 1 : 0;
      ++e;
    }
  }
  m.etags.assign(m.ne, 1);
  m.finalize();
  return m;
}

}  // namespace fim
