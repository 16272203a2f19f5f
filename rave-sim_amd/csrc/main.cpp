// fastwave -- command-line entry point with fast-wave's interface
// (fast-wave/main.cpp:18-57, fast-wave/Readme.md:21-34):
//
//   fastwave <sim_dir> -s <source_idx> [--history_dz <f64>] [-h]
//
// plus options of this build: --dtype c8|c16 (override config.yaml),
// --plain (one kernel per operation), --device <n>, --tune <file>|none.
// Writes <sim_dir>/%08d/detected.npy (float64, (n_steps, n_px)).
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/ravesim.h"

static void usage(const char *argv0) {
  std::printf(
      "Usage: %s <sim_dir> -s <source_idx> [options]\n"
      "  <sim_dir>              Main simulation directory\n"
      "  -s, --source_idx <i>   The index of the source point to simulate (required)\n"
      "  --history_dz <f>       Z-Spacing for history entries (big-wave's history layout: (n_px, entries))\n"
      "  --dtype c8|c16         Override the precision given in config.yaml\n"
      "  --plain                Run one kernel per operation instead of fused passes\n"
      "  --device <n>           HIP device index (default 0)\n"
      "  --tune <file>|none     Tuning table to load (default: data/tuned_mi355x.txt beside\n"
      "                         this executable, when present; entries of other devices are ignored)\n"
      "  -h, --help             print this help screen\n",
      argv0);
}

int main(int argc, char **argv) {
  std::string sim_dir;
  int source_idx = -1;
  bool have_src = false;
  int dtype = 0, mode = RS_MODE_FUSED, device = 0;
  double history_dz = 0;
  std::string tune;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto need = [&](const char *what) -> const char * {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n\n", what);
        usage(argv[0]);
        std::exit(1);
      }
      return argv[++i];
    };
    if (a == "-h" || a == "--help") {
      usage(argv[0]);
      return 0;
    } else if (a == "-s" || a == "--source_idx") {
      source_idx = std::atoi(need("source_idx"));
      have_src = true;
    } else if (a.rfind("--source_idx=", 0) == 0) {
      source_idx = std::atoi(a.c_str() + 13);
      have_src = true;
    } else if (a == "--history_dz") {
      history_dz = std::atof(need("history_dz"));
      if (!(history_dz > 0)) {
        std::fprintf(stderr, "history_dz must be positive\n");
        return 1;
      }
    } else if (a == "--dtype") {
      const std::string d = need("dtype");
      if (d == "c8") dtype = RS_C8;
      else if (d == "c16") dtype = RS_C16;
      else {
        std::fprintf(stderr, "unknown dtype %s\n", d.c_str());
        return 1;
      }
    } else if (a == "--plain") {
      mode = RS_MODE_PLAIN;
    } else if (a == "--device") {
      device = std::atoi(need("device"));
    } else if (a == "--tune") {
      tune = need("tune");
    } else if (!a.empty() && a[0] == '-') {
      std::fprintf(stderr, "unknown option %s\n\n", a.c_str());
      usage(argv[0]);
      return 1;
    } else if (sim_dir.empty()) {
      sim_dir = a;
    } else {
      std::fprintf(stderr, "unexpected argument %s\n\n", a.c_str());
      usage(argv[0]);
      return 1;
    }
  }
  if (tune.empty()) {
    // the table shipped beside the executable (ravesim/data/), if any
    char exe[4096];
    const ssize_t n = readlink("/proc/self/exe", exe, sizeof exe - 1);
    if (n > 0) {
      exe[n] = 0;
      std::string d(exe);
      d = d.substr(0, d.rfind('/') + 1) + "data/tuned_mi355x.txt";
      if (access(d.c_str(), R_OK) == 0) tune = d;
    }
  } else if (tune == "none") {
    tune.clear();
  }
  if (!have_src) {
    std::fprintf(stderr, "Source index is required\n\n");
    usage(argv[0]);
    return 1;
  }
  if (sim_dir.empty()) {
    std::fprintf(stderr, "Simulation directory is required\n\n");
    usage(argv[0]);
    return 1;
  }
  rs_ctx *ctx = nullptr;
  if (rs_ctx_create(device, &ctx) != RS_OK) {
    std::fprintf(stderr, "uncaught exception in the fast-wave framework: %s\n", rs_last_error());
    return 1;
  }
  if (!tune.empty()) {
    int32_t n = 0;
    if (rs_tune_load(ctx, tune.c_str(), &n) != RS_OK) {
      std::fprintf(stderr, "cannot load tuning table %s: %s\n", tune.c_str(), rs_last_error());
      rs_ctx_destroy(ctx);
      return 1;
    }
  }
  rs_run_options o{};
  o.dtype = dtype;
  o.mode = mode;
  o.write_output = 1;
  o.history_dz = history_dz;
  const int rc = rs_run_simdir(ctx, sim_dir.c_str(), source_idx, &o);
  if (rc != RS_OK) std::fprintf(stderr, "uncaught exception in the fast-wave framework: %s\n", rs_last_error());
  rs_ctx_destroy(ctx);
  return rc == RS_OK ? 0 : 1;
}
