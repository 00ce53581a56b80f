// sdpa_amd.h -- host C++ API of the MI355X solver, mirroring class SDPA of
// the reference (sdpa_call.h:91-267): same method names, argument meaning,
// primal/dual convention (the file's primal is SDPA's "reversed" dual) and
// result getters.  The interior-point loop is SDPA::solve
// (sdpa_solve.cpp:59-233); every heavy step runs on the GPU through the C
// ABI in sdpa_gpu.h.  Error behaviour: where the reference calls rError
// (exit), these methods throw std::runtime_error.
//
// A C wrapper (sdpa_solver_*) at the end serves FFI callers (Python ctypes).
#ifndef SDPA_AMD_H
#define SDPA_AMD_H

#include <cstdio>

#ifdef __cplusplus
#include <string>
#include <vector>

#include "sdpa_gpu.h"

namespace sdpa_amd {

struct IterationTrace {
  // printOneIteration's columns in the internal convention (sdpa_io.cpp:622-658)
  double mu, thetaP, thetaD, objP, objD, alphaP, alphaD, beta;
};

class SDPA {
 public:
  enum PhaseType { noINFO, pFEAS, dFEAS, pdFEAS, pdINF, pFEAS_dINF, pINF_dFEAS, pdOPT, pUNBD, dUNBD };
  enum ParameterType { PARAMETER_DEFAULT, PARAMETER_UNSTABLE_BUT_FAST, PARAMETER_STABLE_BUT_SLOW };
  enum ConeType { SDP, SOCP, LP };
  enum SparseType { AUTO, SPARSE, DENSE };

  explicit SDPA(int device = 0);
  ~SDPA();

  void setParameterType(ParameterType type = PARAMETER_DEFAULT);
  void setParameterMaxIteration(int v) { param_.maxIteration = v; }
  void setParameterEpsilonStar(double v) { param_.epsilonStar = v; }
  void setParameterLambdaStar(double v) { param_.lambdaStar = v; }
  void setParameterOmegaStar(double v) { param_.omegaStar = v; }
  void setParameterLowerBound(double v) { param_.lowerBound = v; }
  void setParameterUpperBound(double v) { param_.upperBound = v; }
  void setParameterBetaStar(double v) { param_.betaStar = v; }
  void setParameterBetaBar(double v) { param_.betaBar = v; }
  void setParameterGammaStar(double v) { param_.gammaStar = v; }
  void setParameterEpsilonDash(double v) { param_.epsilonDash = v; }
  // the reference spells this setter "Uppwer" in its examples; both work
  void setParameterUppwerBound(double v) { param_.upperBound = v; }
  // printf formats of printResultXVec / XMat / YMat and of the summary
  // (Parameter::xPrint, XPrint, YPrint, infPrint; "NOPRINT" suppresses)
  void setParameterPrintXVec(const char* f) { xPrint_ = f; }
  void setParameterPrintXMat(const char* f) { XPrint_ = f; }
  void setParameterPrintYMat(const char* f) { YPrint_ = f; }
  void setParameterPrintInformation(const char* f) { infPrint_ = f; }
  const char* getParameterPrintXVec() const { return xPrint_.c_str(); }
  const char* getParameterPrintXMat() const { return XPrint_.c_str(); }
  const char* getParameterPrintYMat() const { return YPrint_.c_str(); }
  const char* getParameterPrintInformation() const { return infPrint_.c_str(); }
  ParameterType getParameterType() const { return paramType_; }
  // Newton::setNumThreads (sdpa_newton.cpp:2127-2149): 0 takes OMP_NUM_THREADS
  // (1 if unset); the count is echoed to the display and the result file.
  // It sized the reference's OpenMP bMat loops; the GPU path keeps it only to
  // report it.
  void setNumThreads(int numThreads = 0);
  int getNumThreads() const { return numThreads_; }
  int getParameterMaxIteration() const { return param_.maxIteration; }
  double getParameterEpsilonStar() const { return param_.epsilonStar; }
  double getParameterLambdaStar() const { return param_.lambdaStar; }
  double getParameterOmegaStar() const { return param_.omegaStar; }
  double getParameterLowerBound() const { return param_.lowerBound; }
  double getParameterUpperBound() const { return param_.upperBound; }
  double getParameterBetaStar() const { return param_.betaStar; }
  double getParameterBetaBar() const { return param_.betaBar; }
  double getParameterGammaStar() const { return param_.gammaStar; }
  double getParameterEpsilonDash() const { return param_.epsilonDash; }
  void setDisplay(FILE* Display = stdout) { display_ = Display; }
  void setResultFile(FILE* fpout = stdout) { fpout_ = fpout; }
  FILE* getDisplay() const { return display_; }
  FILE* getResultFile() const { return fpout_; }
  void setInitPoint(bool isInitPoint) { isInitPoint_ = isInitPoint; }
  bool getInitPoint() const { return isInitPoint_; }
  void setKappa(double kappa) { kappa_ = kappa; }

  void inputConstraintNumber(int m);
  void inputBlockNumber(int nBlock);
  void inputBlockSize(int l, int size);
  void inputBlockType(int l, ConeType coneType);
  void inputCVec(int k, double value);
  void inputElement(int k, int l, int i, int j, double value, bool inputCheck = false);
  void inputInitXVec(int k, double value);
  void inputInitXMat(int l, int i, int j, double value);
  void inputInitYMat(int l, int i, int j, double value);
  void initializeUpperTriangleSpace();
  void initializeUpperTriangle(bool inputTwiceCheck = false);
  void initializeSolve();
  // SDPA::resetParameters (sdpa_solve.cpp:45-56) re-derives mu, the initial
  // residuals, beta, theta, the solve information and the phase from the
  // current parameters.  solve() derives all of them at its start from the
  // parameters set at that time, so a reset between solves has nothing left
  // to do.
  void resetParameters() {}
  void solve();

  double* getResultXVec() { return xVec_.data(); }
  double* getResultXMat(int l);
  double* getResultYMat(int l);
  double getPrimalObj() const { return -objValDual_; }
  double getDualObj() const { return -objValPrimal_; }
  double getPrimalError() const { return normDualMat_; }
  double getDualError() const { return normPrimalVec_; }
  double getDigits() const;
  int getIteration() const { return pIteration_; }
  double getMu() const { return mu_; }
  double getDualityGap() const;
  PhaseType getPhaseValue() const { return (PhaseType)phaseUser_; }
  void getPhaseString(char* str) const;
  double getSolveTime() const { return solveTime_; }
  int getConstraintNumber() const { return m_; }
  int getBlockNumber() const { return nBlock_; }
  int getBlockSize(int l) const;
  ConeType getBlockType(int l) const;

  void readInput(const char* filename, FILE* fpout = nullptr, SparseType type = AUTO);
  void readInit(const char* filename, FILE* fpout = nullptr, SparseType type = AUTO);
  void readParameter(const char* filename, FILE* fpout = nullptr);
  void writeInputSparse(const char* filename, const char* printFormat);
  void writeInitSparse(const char* filename, const char* printFormat);

  // DIMACS errors of the final point, DimacsError[1..6] (index 0 unused)
  void getDimacsError(double* DimacsError);
  void printDimacsError(double* DimacsError, const char* printFormat, FILE* fpout = stdout);
  void printResultXVec(FILE* fp = stdout);
  void printResultXMat(FILE* fp = stdout);
  void printResultYMat(FILE* fp = stdout);
  void printComputationTime(FILE* fp = stdout);
  void printParameters(FILE* fp = stdout);
  static void printSDPAVersion(FILE* fp = stdout);
  void terminate();
  void copyCurrentToInit();

  // additions for this build
  const std::vector<IterationTrace>& trace() const { return trace_; }
  sdpa_gpu_ctx* context() { return ctx_; }
  void loadRecords(int m, int nBlock, const int* blockStruct, const double* c, long nnz,
                   const int* k, const int* l, const int* i, const int* j, const double* v);
  // Start from the given internal-convention point (X, Z packed; y), as
  // setInitPoint(true) + inputInit* would.
  void setInternalInitPoint(const double* X, const double* Z, const double* y);
  double phaseTimes[5] = {0, 0, 0, 0, 0};  // bmat, chol, solve, eig, inv (ms; when timing on)
  // the buckets of the reference's ComputeTime (sdpa_parts.h:48-86) this
  // build measures, seconds of the last solve (device ones when timing on)
  struct ComputeTimes {
    double predictor = 0, corrector = 0, stepPredictor = 0, stepCorrector = 0, mainLoop = 0;
    double makebMat = 0, bF1 = 0, bF3 = 0, cholesky = 0, solve = 0, eigXZ = 0, invXZ = 0;
    double makegVec = 0, makedX = 0, updateRes = 0;
  };
  const ComputeTimes& getComputeTime() const { return computeTime_; }
  void enableTiming(bool on) { timing_ = on; }
  // Extension with no counterpart in the reference: precision 1 assembles
  // the Schur matrix with its F1/F2 products in FP32 and factors it in mixed
  // precision (FP32 factor of the equilibrated matrix on FP32 MFMA, solutions
  // refined to FP64 accuracy against the FP64 matrix) while mu > muSwitch and
  // the refinement converges; from the first stall (or FP32 pivot failure)
  // on, FP64 as the reference.  0 (default): FP64 only.
  void setPrecision(int precision, double muSwitch = 0.0) {
    precision_ = precision;
    muSwitch_ = muSwitch;
  }
  int precision() const { return precision_; }
  double muSwitch() const { return muSwitch_; }
  int failureCode() const { return failure_; }  // 1 chol, 2 step, 3 inverse

 private:
  struct Param {
    int maxIteration;
    double epsilonStar, lambdaStar, omegaStar, lowerBound, upperBound, betaStar, betaBar,
        gammaStar, epsilonDash;
  } param_;
  int device_;
  sdpa_gpu_ctx* ctx_ = nullptr;
  FILE* display_ = nullptr;
  FILE* fpout_ = nullptr;
  int numThreads_ = 1;
  ComputeTimes computeTime_;
  bool isInitPoint_ = false;
  bool timing_ = false;
  int precision_ = 0;
  double muSwitch_ = 0.0;
  ParameterType paramType_ = PARAMETER_DEFAULT;
  std::string xPrint_ = "%+8.3e", XPrint_ = "%+8.3e", YPrint_ = "%+8.3e",
              infPrint_ = "%+10.16e";
  void printParametersWith(FILE* fp, const char* format);
  void printBlocks(FILE* fp, const std::vector<double>& M, const char* format);
  void printLastInfo(FILE* fp);
  double kappa_ = 12.0;
  int m_ = 0, nBlock_ = 0;
  std::vector<int> blockStruct_;
  std::vector<int> coneType_;
  std::vector<double> c_;
  std::vector<int> rk_, rl_, ri_, rj_;
  std::vector<double> rv_;
  bool prepared_ = false;
  long total_ = 0;
  int nDim_ = 0;
  std::vector<long> blockOff_;  // packed offset per user block (SDP n*n, LP entries)
  // initial point (internal convention)
  std::vector<double> initX_, initZ_, initY_;
  // results
  std::vector<double> X_, Z_, y_, xVec_;
  std::vector<IterationTrace> trace_;
  double objValPrimal_ = 0, objValDual_ = 0, normPrimalVec_ = 0, normDualMat_ = 0, mu_ = 0;
  int pIteration_ = 0, phase_ = 0, phaseUser_ = 0, failure_ = 0;
  double solveTime_ = 0;
  void ensureContext();
  void buildLayout();
};

}  // namespace sdpa_amd
extern "C" {
#endif

/* C wrapper for FFI callers.  All return 0 on success, <0 on error. */
typedef struct sdpa_solver sdpa_solver;
sdpa_solver* sdpa_solver_new(int device);
void sdpa_solver_free(sdpa_solver* s);
const char* sdpa_solver_error(sdpa_solver* s);
/* names: maxIteration epsilonStar lambdaStar omegaStar lowerBound upperBound
 * betaStar betaBar gammaStar epsilonDash kappa, and the extensions
 * precision (0 FP64, 1 mixed) and muSwitch (see SDPA::setPrecision) */
int sdpa_solver_set_param(sdpa_solver* s, const char* name, double value);
int sdpa_solver_load(sdpa_solver* s, int m, int nBlock, const int* blockStruct, const double* c,
                     long nnz, const int* k, const int* l, const int* i, const int* j,
                     const double* v);
int sdpa_solver_read_input(sdpa_solver* s, const char* path);
int sdpa_solver_read_parameter(sdpa_solver* s, const char* path);
/* Internal-convention initial point (packed X and Z, y of length m); NULL clears. */
int sdpa_solver_set_init(sdpa_solver* s, const double* X, const double* Z, const double* y);
int sdpa_solver_enable_timing(sdpa_solver* s, int on);
int sdpa_solver_solve(sdpa_solver* s);
/* initializeSolve (upload + structural analysis) and solve, separately: a
 * prepared problem can be solved repeatedly with its data resident. */
int sdpa_solver_prepare(sdpa_solver* s);
int sdpa_solver_run(sdpa_solver* s);
/* out[0..11]: primalObj, dualObj, primalError, dualError, mu, relativeGap,
 * iterations, phase (user convention), failure code, solve seconds, m, packed length */
int sdpa_solver_result(sdpa_solver* s, double* out12);
/* Per-iteration trace, 8 doubles per row (internal convention); returns rows. */
int sdpa_solver_trace(sdpa_solver* s, double* out, int cap_rows);
/* Internal-convention final iterate (X, Z packed; y before the sign flip). */
int sdpa_solver_iterate(sdpa_solver* s, double* X, double* Z, double* y);
/* out[0..4] ms spent in bMat, Cholesky, solves, eigen, inverses (timing on). */
int sdpa_solver_phase_times(sdpa_solver* s, double* out5);
sdpa_gpu_ctx* sdpa_solver_context(sdpa_solver* s);
/* SDPA::readInit / writeInputSparse / writeInitSparse (sdpa_call.cpp:946-990,
 * 1099-1222); the init point read is used when the solve runs with setInitPoint. */
int sdpa_solver_read_init(sdpa_solver* s, const char* path);
int sdpa_solver_write_input(sdpa_solver* s, const char* path, const char* printFormat);
int sdpa_solver_write_init(sdpa_solver* s, const char* path, const char* printFormat);
/* SDPA::getDimacsError (sdpa_call.cpp:849-853): out[1..6] (out[0] = 0). */
int sdpa_solver_dimacs(sdpa_solver* s, double* out7);
/* printResultXVec / XMat / YMat (which = 0, 1, 2) or printComputationTime
 * (which = 3) into the file at path. */
int sdpa_solver_print_result(sdpa_solver* s, int which, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* SDPA_AMD_H */
