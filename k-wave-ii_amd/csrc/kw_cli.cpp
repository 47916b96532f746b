// kspaceFirstOrder-HIP: the command line of the reference's native engines
// (kspaceFirstOrder-OMP / -CUDA) as kspaceFirstOrder3DC.m:468-546 calls it:
//
//   kspaceFirstOrder-HIP -i <input.h5> -o <output.h5> [record flags] [-s start]
//                        [-g device] [-t threads] [-c level] [-r interval]
//                        [--verbose level] [--copy_sensor_mask]
//
// Record flags: --p_raw --p_max --p_min --p_rms --p_max_all --p_min_all
// --p_final --u_raw --u_max --u_min --u_rms --u_max_all --u_min_all --u_final
// --u_non_staggered_raw (none given: --p_raw).  The simulation runs on one
// GPU through kw_run_file (include/kwhip.h).  Exit status 0 on success; on
// failure the message goes to stderr and the status is 1, which the
// reference's MATLAB side reports as an error (kspaceFirstOrder3DC.m:549-567).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "kwhip.h"

namespace {

struct Flag {
  const char* name;
  uint32_t bit;
};
const Flag kFlags[] = {
    {"--p_raw", KW_REC_P},           {"--p_max", KW_REC_P_MAX},
    {"--p_min", KW_REC_P_MIN},       {"--p_rms", KW_REC_P_RMS},
    {"--p_max_all", KW_REC_P_MAX_ALL}, {"--p_min_all", KW_REC_P_MIN_ALL},
    {"--p_final", KW_REC_P_FINAL},   {"--u_raw", KW_REC_U},
    {"--u_max", KW_REC_U_MAX},       {"--u_min", KW_REC_U_MIN},
    {"--u_rms", KW_REC_U_RMS},       {"--u_max_all", KW_REC_U_MAX_ALL},
    {"--u_min_all", KW_REC_U_MIN_ALL}, {"--u_final", KW_REC_U_FINAL},
    {"--u_non_staggered_raw", KW_REC_U_NON_STAGGERED},
};

void usage(FILE* f) {
  std::fprintf(f,
               "Usage: kspaceFirstOrder-HIP -i <input.h5> -o <output.h5> [options]\n"
               "  -s <index>             first time step to record (1-based, default 1)\n"
               "  -g <device>            GPU to run on (default: the first)\n"
               "  -t <threads>, -c <level>, -r <interval>, --verbose <level>,\n"
               "  --copy_sensor_mask     accepted for compatibility (no effect)\n"
               "  record flags           --p_raw --p_max --p_min --p_rms --p_max_all\n"
               "                         --p_min_all --p_final --u_raw --u_max --u_min --u_rms\n"
               "                         --u_max_all --u_min_all --u_final --u_non_staggered_raw\n"
               "                         (none given: --p_raw)\n");
}

bool integer_arg(const char* s, long long& v) {
  char* end = nullptr;
  v = std::strtoll(s, &end, 10);
  return end && *end == '\0' && end != s;
}

}  // namespace

int main(int argc, char** argv) {
  std::string in, out;
  uint32_t rec = 0;
  long long start = 1;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto value = [&](const char* what) -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "Error: %s needs a value\n", what);
        std::exit(1);
      }
      return argv[++i];
    };
    if (a == "-h" || a == "--help") {
      usage(stdout);
      return 0;
    } else if (a == "--version") {
      std::printf("kspaceFirstOrder-HIP (kwhip ABI %d, %s)\n", kw_abi_version(), kw_backend());
      return 0;
    } else if (a == "-i") {
      in = value("-i");
    } else if (a == "-o") {
      out = value("-o");
    } else if (a == "-s") {
      if (!integer_arg(value("-s"), start) || start < 1) {
        std::fprintf(stderr, "Error: -s takes a time step of at least 1\n");
        return 1;
      }
    } else if (a == "-g") {
      long long g;
      if (!integer_arg(value("-g"), g) || g < 0) {
        std::fprintf(stderr, "Error: -g takes a device number\n");
        return 1;
      }
      // restrict the process to that device before the first HIP call
      setenv("HIP_VISIBLE_DEVICES", std::to_string(g).c_str(), 1);
    } else if (a == "-t" || a == "-c" || a == "-r" || a == "--verbose") {
      long long ignored;
      if (!integer_arg(value(a.c_str()), ignored)) {
        std::fprintf(stderr, "Error: %s takes an integer\n", a.c_str());
        return 1;
      }
    } else if (a == "--copy_sensor_mask") {
    } else {
      bool found = false;
      for (const Flag& f : kFlags)
        if (a == f.name) {
          rec |= f.bit;
          found = true;
        }
      if (!found) {
        std::fprintf(stderr, "Error: unknown option '%s' (checkpointing and --benchmark are not "
                             "supported)\n", a.c_str());
        usage(stderr);
        return 1;
      }
    }
  }
  if (in.empty() || out.empty()) {
    std::fprintf(stderr, "Error: the input (-i) and output (-o) files are required\n");
    usage(stderr);
    return 1;
  }
  if (rec == 0) rec = KW_REC_P;
  char err[2048];
  const int rc = kw_run_file(in.c_str(), out.c_str(), rec, (int64_t)start, err, sizeof err);
  if (rc != KW_OK) {
    std::fprintf(stderr, "Error: %s\n", err);
    return 1;
  }
  return 0;
}
