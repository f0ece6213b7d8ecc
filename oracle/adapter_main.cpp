// adapter_main.cpp -- the reference's tw::run (src/main.cpp:84-104) with the
// one line INTEGRATION.md section 2 changes: CufimSolver (the adapter printed
// there, extracted verbatim into oracle/_ref/cufim_adapter.h) in place of
// CudaSolverFIM.  Built by `make -C oracle adapter` against the reference's
// own headers and host code (oracle/_ref/libref_host.so) and libcufim.so;
// test infrastructure only: it shows the adapter compiles, links and runs
// from the reference side.
//
//   adapter_run MESH TAGS CFG PLAN OUT [REPS]
#include <cstdio>
#include <cstdlib>

#include "cufim_adapter.h"
#include "io/io_base.h"
#include "io/io_json.h"
#include "io/vtk.h"

int main(int argc, char **argv) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: adapter_run MESH TAGS CFG PLAN OUT [REPS]\n");
    return 2;
  }
  try {
    tw::mesh_t mesh;
    tw::load_vtk(argv[1], mesh);
    tw::read_vector_txt(argv[2], mesh.atags);
    mesh.compute_base_connectivity(false);
    tw::ek_data ekdata(mesh);
    tw::load_options(argv[3], ekdata.opts);
    tw::load_plan(argv[4], ekdata);
    std::vector<std::vector<tw::dbl_t>> act_ms, apd_ms;
    tw::CufimSolver solver(ekdata);
    const double s = solver.solve(act_ms, apd_ms, argv[5], argc > 6 ? std::atoi(argv[6]) : 1);
    std::printf("Average execution time: %f s\n", s);
  } catch (const std::exception &e) {
    std::printf("adapter error: %s\n", e.what());
    return 1;
  }
  return 0;
}
