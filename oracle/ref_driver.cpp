// ref_driver.cpp -- C entry points over the reference's own HOST code
// (src/core/types.cpp, src/io/vtk.cpp, src/io/io_json.cpp, src/io/io_base.cpp),
// compiled from the sources where they lie into oracle/_ref/libref_host.so by
// `make -C oracle ref`.  Test infrastructure only: it pins the oracle's and
// the product's connectivity, submesh numbering, metric geometry, VTK output
// and config/plan parsing against the reference implementation.
//
// The reference keeps state in C++ objects; this driver holds one mesh and
// one submesh decomposition at a time in file-scope variables and copies
// results into caller-provided buffers (sizes are queried first).
#include "core/types.h"
#include "io/io_base.h"
#include "io/io_json.h"
#include "io/vtk.h"

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {
tw::mesh_t g_mesh;
std::vector<tw::mesh_t> g_subs;
std::vector<std::vector<tw::idx_t>> g_l2g, g_n2a, g_a2a;
std::vector<tw::idx_t> g_g2l;
std::vector<std::map<tw::idx_t, tw::idx_t>> g_a2n;
std::vector<std::map<tw::idx_t, std::vector<tw::idx_t>>> g_a2i;
std::vector<std::vector<tw::ipair_t>> g_n2i;
tw::ek_options g_opts;
std::vector<tw::ek_event> g_events;

template <class T, class U>
void put(const std::vector<T> &v, U *out) {
  for (size_t i = 0; i < v.size(); ++i) out[i] = static_cast<U>(v[i]);
}
}  // namespace

extern "C" {

// load_vtk (src/io/vtk.cpp:11-105) + read_vector_txt of the tags
int ref_load_vtk(const char *vtk, const char *tags, int64_t *nn, int64_t *ne) {
  try {
    g_mesh = tw::mesh_t();
    tw::load_vtk(vtk, g_mesh);
    if (tags && tags[0]) tw::read_vector_txt(tags, g_mesh.atags);
    *nn = (int64_t)(g_mesh.xyz.size() / 3);
    *ne = (int64_t)g_mesh.e2n_cnt.size();
    return 0;
  } catch (...) {
    return -1;
  }
}

// current mesh arrays (e2n as stored: file order until ref_connectivity)
void ref_mesh_get(double *xyz, int64_t *e2n, int32_t *etags) {
  put(g_mesh.xyz, xyz);
  put(g_mesh.e2n_con, e2n);
  if (etags) put(g_mesh.etags, etags);
}

// replaces the mesh by arrays (tets only); fibres (1,0,0) as load_vtk sets
int ref_mesh_set(int64_t nn, int64_t ne, const double *xyz, const int64_t *e2n,
                 const int32_t *atags) {
  g_mesh = tw::mesh_t();
  g_mesh.n_points = nn;
  g_mesh.n_cells = ne;
  g_mesh.xyz.assign(xyz, xyz + 3 * nn);
  g_mesh.e2n_cnt.assign(ne, 4);
  g_mesh.e2n_dsp.resize(ne);
  for (int64_t e = 0; e < ne; ++e) g_mesh.e2n_dsp[e] = 4 * e;
  g_mesh.e2n_con.assign(e2n, e2n + 4 * ne);
  g_mesh.etype.assign(ne, tw::Tetra);
  g_mesh.etags.assign(ne, 1);
  if (atags) g_mesh.atags.assign(atags, atags + ne);
  g_mesh.lon.assign(3 * ne, 0.0);
  for (int64_t e = 0; e < ne; ++e) g_mesh.lon[3 * e] = 1.0;
  return 0;
}

// compute_base_connectivity (src/core/types.cpp:43-225); returns n2n size
int64_t ref_connectivity(void) {
  g_mesh.compute_base_connectivity(false);
  return (int64_t)g_mesh.n2n_con.size();
}

void ref_conn_get(int64_t *e2n_sorted, int32_t *n2e_cnt, int64_t *n2e_dsp,
                  int64_t *n2e_con, int32_t *n2n_cnt, int64_t *n2n_dsp,
                  int64_t *n2n_con) {
  put(g_mesh.e2n_con, e2n_sorted);
  put(g_mesh.n2e_cnt, n2e_cnt);
  put(g_mesh.n2e_dsp, n2e_dsp);
  put(g_mesh.n2e_con, n2e_con);
  put(g_mesh.n2n_cnt, n2n_cnt);
  put(g_mesh.n2n_dsp, n2n_dsp);
  put(g_mesh.n2n_con, n2n_con);
}

// compute_geo (src/core/types.cpp:227-257) on the current (sorted) mesh
void ref_geo(double *geo) {
  g_mesh.compute_geo();
  put(g_mesh.geo, geo);
}

// extract_submeshes2 (src/core/types.cpp:306-570); returns number of submeshes
int64_t ref_extract(void) {
  try {
    return (int64_t)tw::extract_submeshes2(g_mesh, g_mesh.atags, g_subs, g_l2g,
                                           g_g2l, g_n2a, g_a2n, g_a2a, g_a2i, g_n2i);
  } catch (...) {
    return -1;
  }
}

void ref_sub_sizes(int64_t a, int64_t *nn, int64_t *ne, int64_t *nnnc) {
  *nn = (int64_t)g_subs[a].n2e_cnt.size();
  *ne = (int64_t)g_subs[a].e2n_cnt.size();
  *nnnc = (int64_t)g_subs[a].n2n_con.size();
}

void ref_sub_get(int64_t a, int64_t *l2g, int64_t *e2n, double *geo,
                 int32_t *n2e_cnt, int64_t *n2e_con, int32_t *n2n_cnt,
                 int64_t *n2n_con) {
  const tw::mesh_t &s = g_subs[a];
  put(g_l2g[a], l2g);
  put(s.e2n_con, e2n);
  put(s.geo, geo);
  put(s.n2e_cnt, n2e_cnt);
  put(s.n2e_con, n2e_con);
  put(s.n2n_cnt, n2n_cnt);
  put(s.n2n_con, n2n_con);
}

// n2i copies: for each global vertex, (local idx, submesh) pairs
int64_t ref_n2i_size(void) {
  int64_t n = 0;
  for (const auto &v : g_n2i) n += (int64_t)v.size();
  return n;
}

void ref_n2i_get(int64_t *cnt, int64_t *ldx, int64_t *adx) {
  int64_t k = 0;
  for (size_t v = 0; v < g_n2i.size(); ++v) {
    cnt[v] = (int64_t)g_n2i[v].size();
    for (const auto &p : g_n2i[v]) {
      ldx[k] = p.vtx;
      adx[k] = p.adx;
      ++k;
    }
  }
}

// save_vtk (src/io/vtk.cpp:109-212) with point data phi and states
int ref_save_vtk(const char *path, const double *phi, const double *states) {
  try {
    const int64_t nn = (int64_t)(g_mesh.xyz.size() / 3);
    std::vector<std::vector<tw::dbl_t>> data;
    std::vector<std::string> names;
    if (phi) {
      data.emplace_back(phi, phi + nn);
      names.emplace_back("phi");
    }
    if (states) {
      data.emplace_back(states, states + nn);
      names.emplace_back("states");
    }
    tw::save_vtk(path, g_mesh, data, names, {}, {});
    return 0;
  } catch (...) {
    return -1;
  }
}

// load_options (src/io/io_json.cpp:228-296)
int ref_load_options(const char *cfg, int32_t *verbose, int32_t *solvetype,
                     int32_t *rent_en, double *dt, int64_t *n_steps) {
  try {
    g_opts = tw::ek_options();
    tw::load_options(cfg, g_opts);
    *verbose = g_opts.verbose;
    *solvetype = g_opts.solvetype;
    *rent_en = g_opts.rent_en;
    *dt = g_opts.dt;
    *n_steps = (int64_t)g_opts.n_steps;
    return 0;
  } catch (...) {
    return -1;
  }
}

// load_plan (src/io/io_json.cpp:299-318) on a mesh-less ek_data; returns the
// number of events.  dt / n_steps are the plan's overrides.
int ref_load_plan(const char *plan, double *dt, int64_t *n_steps,
                  int64_t *total_vtx) {
  try {
    static tw::mesh_t empty;
    tw::ek_data ek(empty);
    ek.opts = g_opts;
    tw::load_plan(plan, ek);
    g_events = ek.events;
    *dt = ek.opts.dt;
    *n_steps = (int64_t)ek.opts.n_steps;
    int64_t n = 0;
    for (const auto &e : g_events) n += (int64_t)e.vtx.size();
    *total_vtx = n;
    return (int)g_events.size();
  } catch (...) {
    return -1;
  }
}

void ref_plan_events(int32_t *active, double *tstart, double *dur,
                     int64_t *cnt, int64_t *vtx) {
  int64_t k = 0;
  for (size_t i = 0; i < g_events.size(); ++i) {
    active[i] = g_events[i].active;
    tstart[i] = g_events[i].tstart;
    dur[i] = g_events[i].dur;
    cnt[i] = (int64_t)g_events[i].vtx.size();
    for (auto v : g_events[i].vtx) vtx[k++] = v;
  }
}

}  // extern "C"
