// sdpa_main.cpp -- command-line front end with the reference's options
// (sdpa_exe.cpp:30-343): "sdpa DataFile OutputFile [InitialPtFile] [-pt n]
// [-dimacs] [-numThreads n]" or "sdpa [-dd|-ds data] [-id|-is init] -o out
// [-p param] [-pt n] [-k kappa] [-dimacs] [-numThreads n]".  The solve runs on
// the GPU through libsdpa_amd; printing follows the reference's formats.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <chrono>
#include <stdexcept>

#include "../../include/sdpa_amd.h"

using sdpa_amd::SDPA;

static void usage(const char* argv0) {
  std::printf("\n*** Please assign data file and output file.***\n\n");
  std::printf("---- option type 1 ------------\n");
  std::printf("%s DataFile OutputFile [InitialPtFile] [-pt parameters] [-dimacs] [-numThreads numThreads]\n", argv0);
  std::printf("parameters = 0 default, 1 fast (unstable), 2 slow (stable)\n");
  std::printf("  -dimacs : printout dimacs information incurring additional computation cost \n");
  std::printf("  -numThreads: accepted for compatibility (the GPU path ignores it)\n\n");
  std::printf("---- option type 2 ------------\n");
  std::printf("%s [option filename]+ \n", argv0);
  std::printf("  -dd : data dense :: -ds : data sparse     \n");
  std::printf("  -id : init dense :: -is : init sparse     \n");
  std::printf("  -o  : output     :: -p  : parameter       \n");
  std::printf("  -pt : parameters , 0 default, 1 fast (unstable)\n");
  std::printf("                     2 slow (stable)         \n");
  std::printf("  -k  : kappa of the formula choice (default 12)\n");
  std::exit(1);
}

static SDPA::ParameterType ptype(int v) {
  return v == 1 ? SDPA::PARAMETER_UNSTABLE_BUT_FAST
                : (v == 2 ? SDPA::PARAMETER_STABLE_BUT_SLOW : SDPA::PARAMETER_DEFAULT);
}

static bool ends_sparse(const char* f) {
  const size_t n = std::strlen(f);
  return n >= 2 && f[n - 1] == 's' && f[n - 2] == '-';
}

static void now_string(char* out, size_t cap) {
  std::time_t t = std::time(nullptr);
  std::strncpy(out, std::ctime(&t), cap - 1);
  out[cap - 1] = '\0';
  const size_t n = std::strlen(out);
  if (n > 0 && out[n - 1] == '\n') out[n - 1] = '\0';
}

int main(int argc, char** argv) {
  const auto t0 = std::chrono::steady_clock::now();
  char when[128];
  now_string(when, sizeof(when));
  std::printf("SDPA start at [%s]\n", when);
  if (argc == 1) usage(argv[0]);
  std::setbuf(stdout, nullptr);

  const char *input = nullptr, *result = nullptr, *init = nullptr, *param = nullptr;
  SDPA::SparseType inSparse = SDPA::DENSE, initSparse = SDPA::DENSE;
  SDPA::ParameterType type = SDPA::PARAMETER_DEFAULT;
  bool dimacs = false;
  double kappa = 12.0;
  bool kappaSet = false;
  if (argv[1][0] == '-') {
    for (int a = 1; a < argc; ++a) {
      const char* o = argv[a];
      const bool hasArg = a + 1 < argc;
      if (!std::strcmp(o, "-dd") && hasArg) { input = argv[++a]; inSparse = SDPA::DENSE; }
      else if (!std::strcmp(o, "-ds") && hasArg) { input = argv[++a]; inSparse = SDPA::SPARSE; }
      else if (!std::strcmp(o, "-id") && hasArg) { init = argv[++a]; initSparse = SDPA::DENSE; }
      else if (!std::strcmp(o, "-is") && hasArg) { init = argv[++a]; initSparse = SDPA::SPARSE; }
      else if (!std::strcmp(o, "-o") && hasArg) result = argv[++a];
      else if (!std::strcmp(o, "-p") && hasArg) param = argv[++a];
      else if (!std::strcmp(o, "-k") && hasArg) { kappa = std::atof(argv[++a]); kappaSet = true; }
      else if (!std::strcmp(o, "-dimacs")) dimacs = true;
      else if (!std::strcmp(o, "-pt") && hasArg) { type = ptype(std::atoi(argv[++a])); param = nullptr; }
      else if (!std::strcmp(o, "-numThreads") && hasArg) ++a;
    }
  } else {
    if (argc < 3) usage(argv[0]);
    input = argv[1];
    if (ends_sparse(input)) inSparse = SDPA::SPARSE;
    result = argv[2];
    param = "./param.sdpa";
    for (int a = 3; a < argc; ++a) {
      if (!std::strcmp(argv[a], "-dimacs")) dimacs = true;
      else if (!std::strcmp(argv[a], "-numThreads") && a + 1 < argc) ++a;
      else if (!std::strcmp(argv[a], "-pt") && a + 1 < argc) { type = ptype(std::atoi(argv[++a])); param = nullptr; }
      else {
        init = argv[a];
        if (ends_sparse(init)) initSparse = SDPA::SPARSE;
      }
    }
  }
  if (!input || !result) usage(argv[0]);

  try {
    SDPA P;
    P.setDisplay(stdout);
    if (kappaSet) P.setKappa(kappa);
    FILE* out = std::fopen(result, "w");
    if (!out) { std::fprintf(stderr, "Cannot Open Result File : %s\n", result); return 1; }
    std::fprintf(out, "SDPA start at [%s]\n", when);
    P.setResultFile(out);
    if (!param) {
      const char* names[3] = {"DEFAULT", "UNSTABLE_BUT_FAST", "STABLE_BUT_SLOW"};
      const int ti = type == SDPA::PARAMETER_UNSTABLE_BUT_FAST ? 1 : (type == SDPA::PARAMETER_STABLE_BUT_SLOW ? 2 : 0);
      std::printf("set   is %s\n", names[ti]);
      std::fprintf(out, "set   is %s\n", names[ti]);
      P.setParameterType(type);
    } else {
      std::printf("param is %s\n", param);
      P.readParameter(param, out);
    }
    std::printf("data  is %s : %s\n", input, inSparse == SDPA::SPARSE ? "sparse" : "dense");
    P.readInput(input, out, inSparse);
    if (init) {
      P.setInitPoint(true);
      std::printf("init  is %s : %s\n", init, initSparse == SDPA::SPARSE ? "sparse" : "dense");
      P.readInit(init, out, initSparse);
    }
    std::printf("out   is %s\n", result);
    std::fprintf(out, "out    is %s\n", result);
    if (dimacs) {
      std::printf("Dimacs information will be computed after the iteration.\n");
      std::fprintf(out, "Dimacs information will be computed after the iteration.\n");
    }
    P.initializeSolve();
    P.solve();
    if (dimacs) {
      double e[7];
      P.getDimacsError(e);
      P.printDimacsError(e, P.getParameterPrintInformation(), stdout);
      P.printDimacsError(e, P.getParameterPrintInformation(), out);
    }
    P.terminate();
    now_string(when, sizeof(when));
    const double all = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("SDPA end at [%s]\n", when);
    std::fprintf(out, "SDPA end at [%s]\n", when);
    std::printf("ALL TIME = %.6lf\n", all);
    std::fprintf(out, "ALL TIME = %.6lf\n", all);
    std::fclose(out);
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "%s\n", ex.what());
    return 1;
  }
  return 0;
}
