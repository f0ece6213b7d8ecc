// host_capi.cpp -- C ABI of the host-side objects (fim_mesh, fim_pack) and
// the thread-local error message.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "host.hpp"
#include "solver_fim.hpp"

namespace fim {
namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F &&f) {
  try {
    f();
    return FIM_OK;
  } catch (const Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "out of host memory";
    return FIM_ELIMIT;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return FIM_EINVAL;
  } catch (...) {
    g_last_error = "unknown error";
    return FIM_EINVAL;
  }
}
}  // namespace

void set_last_error(const std::string &msg) { g_last_error = msg; }

}  // namespace fim

using namespace fim;

extern "C" {

const char *fim_last_error(void) { return g_last_error.c_str(); }

int fim_mesh_create(int64_t nn, int64_t ne, const double *xyz, const int64_t *e2n,
                    const double *fib, const double *vel, const int32_t *etags, fim_mesh **out) {
  return guarded([&] {
    if (!out || !xyz || !e2n || nn <= 0 || ne <= 0) throw Error(FIM_EINVAL, "bad mesh arguments");
    *out = nullptr;
    fim_mesh *m = new fim_mesh();
    try {
      m->m.nn = nn;
      m->m.ne = ne;
      m->m.xyz.assign(xyz, xyz + 3 * nn);
      m->m.e2n.assign(e2n, e2n + 4 * ne);
      if (fib) m->m.fib.assign(fib, fib + 3 * ne);
      if (vel) m->m.vel.assign(vel, vel + 2 * ne);
      if (etags) m->m.etags.assign(etags, etags + ne);
      m->m.finalize();
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int fim_mesh_load_vtk(const char *path, fim_mesh **out) {
  return guarded([&] {
    if (!path || !out) throw Error(FIM_EINVAL, "null argument");
    *out = nullptr;
    fim_mesh *m = new fim_mesh();
    try {
      m->m = load_vtk(path);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int fim_mesh_read_tags(fim_mesh *m, const char *path) {
  return guarded([&] {
    if (!m || !path) throw Error(FIM_EINVAL, "null argument");
    std::vector<int32_t> t = read_tags(path);
    if ((int64_t)t.size() != m->m.ne)
      throw Error(FIM_EINVAL, "tags file has " + std::to_string(t.size()) + " entries, mesh has " +
                                  std::to_string(m->m.ne) + " elements");
    m->m.atags = std::move(t);
  });
}

int fim_mesh_set_tags(fim_mesh *m, const int32_t *atags) {
  return guarded([&] {
    if (!m || !atags) throw Error(FIM_EINVAL, "null argument");
    m->m.atags.assign(atags, atags + m->m.ne);
  });
}

int fim_mesh_set_regions(fim_mesh *m, const uint8_t *region) {
  return guarded([&] {
    if (!m || !region) throw Error(FIM_EINVAL, "null argument");
    for (int64_t e = 0; e < m->m.ne; ++e)
      if (region[e] > 3) throw Error(FIM_EINVAL, "region ids must be 0..3");
    m->m.region.assign(region, region + m->m.ne);
  });
}

int fim_mesh_info(const fim_mesh *m, int64_t *nn, int64_t *ne) {
  return guarded([&] {
    if (!m) throw Error(FIM_EINVAL, "null mesh");
    if (nn) *nn = m->m.nn;
    if (ne) *ne = m->m.ne;
  });
}

int fim_mesh_get(const fim_mesh *m, double *xyz, int64_t *e2n, int32_t *etags) {
  return guarded([&] {
    if (!m) throw Error(FIM_EINVAL, "null mesh");
    if (xyz) std::memcpy(xyz, m->m.xyz.data(), sizeof(double) * m->m.xyz.size());
    if (e2n) std::memcpy(e2n, m->m.e2n.data(), sizeof(int64_t) * m->m.e2n.size());
    if (etags) std::memcpy(etags, m->m.etags.data(), sizeof(int32_t) * m->m.etags.size());
  });
}

int fim_mesh_get_material(const fim_mesh *m, double *fib, double *vel) {
  return guarded([&] {
    if (!m) throw Error(FIM_EINVAL, "null mesh");
    if (fib) std::memcpy(fib, m->m.fib.data(), sizeof(double) * m->m.fib.size());
    if (vel) std::memcpy(vel, m->m.vel.data(), sizeof(double) * m->m.vel.size());
  });
}

double fim_mesh_front_width(const fim_mesh *m) { return m ? front_width(m->m) : -1.0; }

int fim_mesh_save_vtk(const fim_mesh *m, const char *path, const double *phi, const int32_t *state) {
  return guarded([&] {
    if (!m || !path) throw Error(FIM_EINVAL, "null argument");
    save_vtk(m->m, path, phi, state);
  });
}

void fim_mesh_destroy(fim_mesh *m) { delete m; }

int fim_mesh_gen_shell(double ao, double bo, double ai, double bi, double h, int32_t n_scars,
                       double r0, double r1, uint64_t seed, double vf, double vc, double vs,
                       fim_mesh **out) {
  return guarded([&] {
    if (!out) throw Error(FIM_EINVAL, "null argument");
    *out = nullptr;
    fim_mesh *m = new fim_mesh();
    try {
      m->m = gen_shell(ao, bo, ai, bi, h, n_scars, r0, r1, seed, vf, vc, vs);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int fim_mesh_gen_box(double lx, double ly, double lz, int nx, int ny, int nz, double jitter,
                     uint64_t seed, double angle0, double angle1, double vf, double vc,
                     fim_mesh **out) {
  return guarded([&] {
    if (!out) throw Error(FIM_EINVAL, "null argument");
    *out = nullptr;
    fim_mesh *m = new fim_mesh();
    try {
      m->m = gen_box(lx, ly, lz, nx, ny, nz, jitter, seed, angle0, angle1, vf, vc);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int fim_mesh_concat(const fim_mesh *const *parts, int32_t n, fim_mesh **out) {
  return guarded([&] {
    if (!parts || n <= 0 || !out) throw Error(FIM_EINVAL, "bad concat arguments");
    *out = nullptr;
    fim_mesh *u = new fim_mesh();
    try {
      Mesh &m = u->m;
      bool any_region = false, all_tags = true;
      for (int32_t i = 0; i < n; ++i) {
        if (!parts[i]) throw Error(FIM_EINVAL, "null mesh in concat");
        any_region |= !parts[i]->m.region.empty();
        all_tags &= !parts[i]->m.atags.empty();
      }
      int32_t tag_base = 0;
      for (int32_t i = 0; i < n; ++i) {
        const Mesh &p = parts[i]->m;
        const int64_t vo = m.nn;
        m.xyz.insert(m.xyz.end(), p.xyz.begin(), p.xyz.end());
        for (int64_t v : p.e2n) m.e2n.push_back(v + vo);
        m.fib.insert(m.fib.end(), p.fib.begin(), p.fib.end());
        m.vel.insert(m.vel.end(), p.vel.begin(), p.vel.end());
        m.etags.insert(m.etags.end(), p.etags.begin(), p.etags.end());
        if (any_region) {
          if (p.region.empty())
            m.region.insert(m.region.end(), p.ne, 0);
          else
            m.region.insert(m.region.end(), p.region.begin(), p.region.end());
        }
        if (all_tags) {  // keep the parts' partitions apart
          int32_t mx = 0;
          for (int32_t t : p.atags) {
            m.atags.push_back(t + tag_base);
            mx = std::max(mx, t);
          }
          tag_base += mx + 1;
        }
        m.nn += p.nn;
        m.ne += p.ne;
      }
      m.finalize();
    } catch (...) {
      delete u;
      throw;
    }
    *out = u;
  });
}

int fim_pack_build(const fim_mesh *m, const fim_part_opts *o, fim_pack **out) {
  return guarded([&] {
    if (!m || !o || !out) throw Error(FIM_EINVAL, "null argument");
    *out = nullptr;
    Partition part;
    if (o->source == FIM_PART_TAGS) {
      if (m->m.atags.empty()) throw Error(FIM_EINVAL, "mesh has no partition tags");
      part = partition_from_tags(m->m, m->m.atags);
    } else if (o->source == FIM_PART_RCB) {
      part = partition_rcb(m->m, o->max_nn > 0 ? o->max_nn : 512, o->max_ne, o->seed);
    } else {
      throw Error(FIM_EINVAL, "unknown partition source");
    }
    fim_pack *p = new fim_pack();
    try {
      p->p = build_pack(m->m, part, o->source);
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

int fim_pack_info(const fim_pack *p, fim_pack_info_t *info) {
  return guarded([&] {
    if (!p || !info) throw Error(FIM_EINVAL, "null argument");
    const Pack &k = p->p;
    info->nn_global = k.nn_global;
    info->ne_global = k.ne_global;
    info->nsub = k.nsub;
    info->ncopy = k.ncopy;
    info->n_iface_copies = (int64_t)k.if_l.size();
    info->nn_max = k.nn_max;
    info->ne_max = k.ne_max;
    info->device_bytes_mesh = 32 * k.ne_global + 4 * k.ncopy;
    info->lds_bytes = converge_lds_bytes(k.nn_max, k.ne_max);
  });
}

int fim_pack_partition(const fim_pack *p, int64_t *e2s, int64_t *c2g) {
  return guarded([&] {
    if (!p) throw Error(FIM_EINVAL, "null pack");
    if (e2s) std::memcpy(e2s, p->p.part.e2s.data(), sizeof(int64_t) * p->p.part.e2s.size());
    if (c2g) std::memcpy(c2g, p->p.part.l2g.data(), sizeof(int64_t) * p->p.part.l2g.size());
  });
}

void fim_pack_destroy(fim_pack *p) { delete p; }

int fim_pack_save_tags(const fim_pack *p, const char *path) {
  return guarded([&] {
    if (!p || !path) throw Error(FIM_EINVAL, "null argument");
    FILE *f = std::fopen(path, "w");
    if (!f) throw Error(FIM_EIO, std::string("cannot write ") + path);
    for (int64_t s : p->p.part.e2s) std::fprintf(f, "%lld\n", (long long)s);
    if (std::fclose(f) != 0) throw Error(FIM_EIO, std::string("cannot write ") + path);
  });
}

int fim_mesh_apply_plan_speeds(fim_mesh *m, const char *plan_path, int64_t *n_set) {
  return guarded([&] {
    if (!m || !plan_path) throw Error(FIM_EINVAL, "null argument");
    Options o;
    const Plan plan = load_plan(plan_path, o);
    Mesh &mm = m->m;
    int64_t cnt = 0;
    for (int64_t e = 0; e < mm.ne; ++e) {
      const int32_t tag = mm.etags.empty() ? 0 : mm.etags[e];
      for (const Region &r : plan.regions)
        if (std::find(r.tags.begin(), r.tags.end(), tag) != r.tags.end()) {
          if (!(r.vf > 0) || !(r.vs > 0)) throw Error(FIM_EINVAL, "plan speeds must be positive");
          mm.vel[2 * e] = r.vf;
          mm.vel[2 * e + 1] = r.vs;
          ++cnt;
          break;
        }
    }
    if (n_set) *n_set = cnt;
  });
}

int fim_config_load_options(const char *cfg_path, fim_options_t *out) {
  return guarded([&] {
    if (!cfg_path || !out) throw Error(FIM_EINVAL, "null argument");
    const Options o = load_options(cfg_path);
    *out = fim_options_t{o.verbose, o.method_id, o.solvetype, o.rent_en, o.dt, o.n_steps};
  });
}

int fim_config_load_plan(const char *plan_path, fim_options_t *opts, int32_t *n_events,
                         int64_t *n_vtx) {
  return guarded([&] {
    if (!plan_path || !opts) throw Error(FIM_EINVAL, "null argument");
    Options o;
    o.dt = opts->dt;
    o.n_steps = opts->n_steps;
    const Plan p = load_plan(plan_path, o);
    opts->dt = o.dt;
    opts->n_steps = o.n_steps;
    int64_t nv = 0;
    for (const auto &e : p.events) nv += (int64_t)e.vtx.size();
    if (n_events) *n_events = (int32_t)p.events.size();
    if (n_vtx) *n_vtx = nv;
  });
}

int fim_config_plan_events(const char *plan_path, int32_t *active, double *tstart, double *dur,
                           int64_t *cnt, int64_t *vtx) {
  return guarded([&] {
    if (!plan_path) throw Error(FIM_EINVAL, "null argument");
    Options o;
    const Plan p = load_plan(plan_path, o);
    int64_t k = 0;
    for (size_t i = 0; i < p.events.size(); ++i) {
      if (active) active[i] = p.events[i].active;
      if (tstart) tstart[i] = p.events[i].tstart;
      if (dur) dur[i] = p.events[i].dur;
      if (cnt) cnt[i] = (int64_t)p.events[i].vtx.size();
      for (int64_t v : p.events[i].vtx) {
        if (vtx) vtx[k] = v;
        ++k;
      }
    }
  });
}

int fim_run_reference(const char *mesh_path, const char *part_path, const char *plan_path,
                      const char *out_path, int32_t num_reps, const char *cfg_path,
                      int32_t device, double *avg_s) {
  return guarded([&] {
    if (!mesh_path || !part_path || !plan_path) throw Error(FIM_EINVAL, "null argument");
    const double s = run(mesh_path, part_path, plan_path, out_path ? out_path : "", num_reps,
                         cfg_path ? cfg_path : "data/default.cfg", device);
    if (avg_s) *avg_s = s;
  });
}

}  // extern "C"
