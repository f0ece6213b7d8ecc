// cli_main.cpp -- `cufim`, the reference-format command line
// (src/main.cpp:170-237):
//   cufim --mesh=M.vtk --part=M.tags --plan=M.plan.json --odir=out.vtk [--reps=N]
//   cufim --info | --test | --help
// Reads data/default.cfg from the working directory as the reference does
// (override with --cfg=PATH); --device=N selects the GPU.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <unistd.h>

#include "cufim.h"
#include "host.hpp"
#include "solver_fim.hpp"

namespace {

void usage() {
  std::printf(
      "Usage: cufim [--help] [--mesh STR] [--part STR] [--plan STR] [--odir STR]\n"
      "             [--info] [--test] [--reps INT] [--num_threads INT] [--cfg STR] [--device INT]\n"
      "             [--timed-events]\n\n"
      "FIM activation-time solver for tetrahedral meshes on one MI355X.\n\n"
      "  -m, --mesh         mesh file (legacy ASCII VTK)\n"
      "  -p, --part         partitioning file (.tags, one integer per element)\n"
      "  -s, --plan         plan file (JSON)\n"
      "  -o, --odir         output file (legacy VTK with phi and states)\n"
      "  -i, --info         outputs gpu information\n"
      "  -t, --test         run the unit test configuration (ut_mesh)\n"
      "  -r, --reps         number of runs/repetitions [default: 1]\n"
      "  -np, --num_threads accepted for compatibility (unused) [default: 8]\n"
      "  --cfg              options file [default: data/default.cfg]\n"
      "  --device           GPU ordinal [default: 0]\n"
      "  --timed-events     fire each event in the time window of its start time (multi-beat\n"
      "                     pacing); by default every event fires at t = 0, as in cufim-nvlink\n");
}

// directory of this executable's repository tests/golden/ut_mesh, for --test
std::string golden_dir() {
  char buf[4096];
  const ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
  if (n <= 0) return "tests/golden/ut_mesh";
  buf[n] = 0;
  std::string p(buf);
  for (int up = 0; up < 3; ++up) p = p.substr(0, p.find_last_of('/'));  // bin, pkg, repo
  return p + "/tests/golden/ut_mesh";
}

}  // namespace

int main(int argc, char **argv) {
  static const std::map<std::string, std::string> alias = {
      {"-m", "--mesh"}, {"-p", "--part"}, {"-s", "--plan"}, {"-o", "--odir"},
      {"-i", "--info"}, {"-t", "--test"}, {"-r", "--reps"}, {"-np", "--num_threads"},
      {"-h", "--help"}};
  std::map<std::string, std::string> args;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i], key = a, val;
    const size_t eq = a.find('=');
    if (eq != std::string::npos) {
      key = a.substr(0, eq);
      val = a.substr(eq + 1);
    }
    auto al = alias.find(key);
    if (al != alias.end()) key = al->second;
    if (key != "--mesh" && key != "--part" && key != "--plan" && key != "--odir" &&
        key != "--info" && key != "--test" && key != "--reps" && key != "--num_threads" &&
        key != "--help" && key != "--cfg" && key != "--device" && key != "--timed-events") {
      std::fprintf(stderr, "Unknown argument: %s\n", a.c_str());
      usage();
      return EXIT_FAILURE;
    }
    if (eq == std::string::npos && key != "--info" && key != "--test" && key != "--help" &&
        key != "--timed-events") {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "%s: expected 1 argument\n", key.c_str());
        return EXIT_FAILURE;
      }
      val = argv[++i];
    }
    args[key] = val;
  }
  int code = EXIT_SUCCESS;
  try {
    const int reps = args.count("--reps") ? std::stoi(args["--reps"]) : 1;
    const int device = args.count("--device") ? std::stoi(args["--device"]) : 0;
    const std::string cfg = args.count("--cfg") ? args["--cfg"] : "data/default.cfg";
    if (args.count("--help")) {
      usage();
    } else if (args.count("--info")) {
      char buf[1024];
      if (fim_device_info(device, buf, sizeof(buf)) != FIM_OK) throw fim::Error(FIM_EDEVICE, fim_last_error());
      std::printf("%s\n", buf);
    } else if (args.count("--test")) {
      const std::string d = golden_dir();
      fim::run(d + "/ut_mesh.vtk", d + "/ut_mesh.tags", d + "/ut_mesh.plan.json",
               args.count("--odir") ? args["--odir"] : "cufim-ut_mesh-out.vtk", 1, cfg, device);
    } else if (reps < 1) {
      std::printf("invalid number of reps!\n");
    } else if (args.count("--mesh") && args.count("--part") && args.count("--plan") &&
               args.count("--odir")) {
      const double s = fim::run(args["--mesh"], args["--part"], args["--plan"], args["--odir"],
                                reps, cfg, device, args.count("--timed-events") > 0);
      std::printf("Average execution time for %d runs: %f s\n", reps, s);
    } else {
      usage();
    }
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    code = EXIT_FAILURE;
  }
  return code;
}
