// solver_fim.hpp -- C++ host interface mirroring the reference solver class
// and its configuration loaders, built on the C ABI (include/cufim.h).
//
//   reference                                   here
//   tw::ek_options / load_options               fim::Options / load_options
//     (src/core/types.h:258-293, src/io/io_json.cpp:228-296)
//   tw::ek_event / load_plan                    fim::Event, fim::Plan / load_plan
//     (src/core/types.h:112-130, src/io/io_json.cpp:82-223, 299-318)
//   tw::CudaSolverFIM(ek_data) + solve()        fim::SolverFIM + solve()
//     (src/eikonal/cuda_fim.h:1443-1518, cuda_fim.cu:1218-1501, 2386-2503)
//   tw::run(mesh, part, plan, out, reps)        fim::run
//     (src/main.cpp:84-104)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "cufim.h"

namespace fim {

// data/default.cfg: nine positional lines (src/io/io_json.cpp:228-296)
struct Options {
  int verbose = 0;
  int method_id = 1;   // the reference forces FIM
  int solvetype = 1;   // 0 lin, 1 vol (parsed, not used by the solver)
  int rent_en = 1;
  double dt = 1000.0;  // us (the file holds ms)
  int64_t n_steps = 0;
};

struct Event {
  int id = 0, active = 0;
  double tstart = 0.0, dur = 0.0;  // us
  std::vector<int64_t> vtx;
};

struct Region {
  std::string func;
  std::vector<int32_t> tags;
  int active = 1;
  double vf = 0.6, vs = 0.6, vn = 0.6;  // m/s = um/us
  double apd = 3e5;                     // us
};

struct Plan {
  std::vector<Event> events;
  std::vector<Region> regions;
};

// throws fim::Error(FIM_EIO / FIM_EINVAL) with the reference's messages
Options load_options(const std::string &cfg_path);
// the plan overrides dt and n_steps (src/io/io_json.cpp:145-146)
Plan load_plan(const std::string &plan_path, Options &opts);

class SolverFIM {
 public:
  // Partition tags must be set on the mesh (fim_mesh_read_tags); the mesh
  // must outlive the solver, as ek_data's mesh reference does.
  SolverFIM(fim_mesh *mesh, const Options &opts, const std::vector<Event> &events,
            int device = 0);
  ~SolverFIM();
  SolverFIM(const SolverFIM &) = delete;
  SolverFIM &operator=(const SolverFIM &) = delete;

  // Runs num_reps repetitions of reset, stimulus seeding at t = 0 and
  // n_steps time windows; writes the merged phi/states to `file` (legacy
  // VTK) unless it is empty; returns the mean step-loop time in seconds.
  // act_ms / apd_ms are accepted for interface parity and left untouched,
  // as in the reference (cuda_fim.cu:2492).
  double solve(std::vector<std::vector<double>> &act_ms,
               std::vector<std::vector<double>> &apd_ms, const std::string &file,
               int num_reps);
  double operator()(std::vector<std::vector<double>> &act_ms,
                    std::vector<std::vector<double>> &apd_ms, const std::string &file,
                    int num_reps) {
    return solve(act_ms, apd_ms, file, num_reps);
  }

  // Timed events (off by default, as in the reference): event i fires in
  // the time window that holds its start time, with phi = tstart, and every
  // active event fires -- multi-beat pacing with re-entry through the
  // tail -> far states.  The reference fires every event at t = 0
  // (cuda_fim.cu:632-683).
  void set_timed_events(bool on) { timed_events_ = on; }

  const std::vector<double> &phi() const { return phi_; }
  const std::vector<int32_t> &states() const { return states_; }
  double front_width() const { return front_width_; }

 private:
  fim_mesh *mesh_;
  Options opts_;
  bool timed_events_ = false;
  std::vector<Event> events_;
  fim_pack *pack_ = nullptr;
  fim_ctx *ctx_ = nullptr;
  double front_width_ = 0.0;
  std::vector<double> phi_;
  std::vector<int32_t> states_;
};

// load mesh + tags, data/default.cfg (relative to the working directory, as
// the reference does), plan; solve; write `out_path`.  Returns mean seconds.
double run(const std::string &mesh_path, const std::string &part_path,
           const std::string &plan_path, const std::string &out_path, int num_reps,
           const std::string &cfg_path = "data/default.cfg", int device = 0,
           bool timed_events = false);

}  // namespace fim
