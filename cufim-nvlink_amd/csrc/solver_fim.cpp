// solver_fim.cpp -- fim::SolverFIM (host mirror of the reference solver
// class) and fim::run, on the C ABI.
#include "solver_fim.hpp"

#include <chrono>
#include <cstdio>

#include "host.hpp"

namespace fim {

namespace {
void ck(int rc) {
  if (rc != FIM_OK) throw Error(rc, fim_last_error());
}
}  // namespace

// setup: partition by the mesh's tags (reference numbering), pack, upload
// (src/eikonal/cuda_fim.cu:1218-1501)
SolverFIM::SolverFIM(fim_mesh *mesh, const Options &opts, const std::vector<Event> &events,
                     int device)
    : mesh_(mesh), opts_(opts), events_(events) {
  if (!mesh_) throw Error(FIM_EINVAL, "SolverFIM: null mesh");
  fim_part_opts po{FIM_PART_TAGS, 0, 0, 0};
  ck(fim_pack_build(mesh_, &po, &pack_));
  fim_opts o{};
  o.mode = FIM_MODE_COMPAT;
  o.local = FIM_LOCAL_REF;
  o.n_instances = 1;
  o.device = device;
  o.block = 256;
  o.apd = 3e5;
  const int rc = fim_create(pack_, &o, &ctx_);
  if (rc != FIM_OK) {
    const std::string msg = fim_last_error();
    fim_pack_destroy(pack_);
    pack_ = nullptr;
    throw Error(rc, msg);
  }
  front_width_ = fim_mesh_front_width(mesh_);
  if (opts_.verbose) std::printf("DWF: %g\n", front_width_);
}

SolverFIM::~SolverFIM() {
  fim_destroy(ctx_);
  fim_pack_destroy(pack_);
}

// src/eikonal/cuda_fim.cu:2386-2503
double SolverFIM::solve(std::vector<std::vector<double>> &, std::vector<std::vector<double>> &,
                        const std::string &file, int num_reps) {
  if (num_reps < 1) throw Error(FIM_EINVAL, "invalid number of reps!");
  double sum_s = 0.0;
  for (int rid = 0; rid < num_reps; ++rid) {
    ck(fim_reset(ctx_));
    // events fire at cur_time 0; the scan stops at the first event that is
    // inactive or starts before 0 (cuda_fim.cu:636-640)
    if (!timed_events_)
      for (const Event &ev : events_) {
        if (!(ev.tstart >= 0.0 && ev.active)) break;
        ck(fim_seed(ctx_, 0, ev.vtx.data(), (int64_t)ev.vtx.size(), 0.0));
      }
    double cur = 0.0;
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t step = 0; step < opts_.n_steps; ++step) {
      const double phi_start = cur;
      cur += opts_.dt;
      if (timed_events_)
        for (const Event &ev : events_)
          if (ev.active && ev.tstart >= phi_start && ev.tstart < cur)
            ck(fim_seed(ctx_, 0, ev.vtx.data(), (int64_t)ev.vtx.size(), ev.tstart));
      ck(fim_step(ctx_, phi_start, cur, opts_.dt, front_width_, nullptr));
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    sum_s += s;
    std::printf("run %2d took %.5f s.\n", rid + 1, s);
  }
  int64_t nn = 0;
  ck(fim_mesh_info(mesh_, &nn, nullptr));
  phi_.assign(nn, 0.0);
  states_.assign(nn, FIM_FAR);
  ck(fim_download(ctx_, 0, phi_.data(), states_.data()));
  if (!file.empty()) ck(fim_mesh_save_vtk(mesh_, file.c_str(), phi_.data(), states_.data()));
  return sum_s / num_reps;
}

// src/main.cpp:84-104
double run(const std::string &mesh_path, const std::string &part_path,
           const std::string &plan_path, const std::string &out_path, int num_reps,
           const std::string &cfg_path, int device, bool timed_events) {
  fim_mesh *m = nullptr;
  ck(fim_mesh_load_vtk(mesh_path.c_str(), &m));
  struct Guard {
    fim_mesh *m;
    ~Guard() { fim_mesh_destroy(m); }
  } g{m};
  ck(fim_mesh_read_tags(m, part_path.c_str()));
  Options opts = load_options(cfg_path);
  Plan plan = load_plan(plan_path, opts);
  std::vector<std::vector<double>> act, apd;
  SolverFIM solver(m, opts, plan.events, device);
  solver.set_timed_events(timed_events);
  return solver.solve(act, apd, out_path, num_reps);
}

}  // namespace fim
