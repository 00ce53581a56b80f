// sdpa_host.cpp -- the host interior-point driver (class sdpa_amd::SDPA) and
// its C wrapper.  The loop restates SDPA::solve (sdpa_solve.cpp:59-233) and
// the control scalars of sdpa_parts.cpp (StepLength 435-625,
// DirectionParameter 653-706, Switch 733-743, AverageComplementarity 770-790,
// RatioInitResCurrentRes 822-866, SolveInfo 914-1060, Phase 1150-1242); every
// matrix-sized step is a call into the C ABI of sdpa_gpu.h.
#include "../../include/sdpa_amd.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace sdpa_amd {

namespace {

enum { D_CX = 0, D_BY, D_CDX, D_BDY, D_XDZ, D_DXZ, D_DXDZ, D_XZ, D_TRX, D_TRZ, D_X0Z, D_XZ0 };

void check(int rc, sdpa_gpu_ctx* ctx, const char* what) {
  if (rc < 0) {
    std::string msg = std::string("sdpa_amd: ") + what + " failed (" + std::to_string(rc) + ")";
    if (ctx) msg += std::string(": ") + sdpa_gpu_last_error(ctx);
    throw std::runtime_error(msg);
  }
}

// Phase::reverse (sdpa_parts.cpp:1244-1264)
int reverse_phase(int p) {
  switch (p) {
    case SDPA::pFEAS: return SDPA::dFEAS;
    case SDPA::dFEAS: return SDPA::pFEAS;
    case SDPA::pFEAS_dINF: return SDPA::pINF_dFEAS;
    case SDPA::pINF_dFEAS: return SDPA::pFEAS_dINF;
    case SDPA::pUNBD: return SDPA::dUNBD;
    case SDPA::dUNBD: return SDPA::pUNBD;
    default: return p;
  }
}

// the number scan of IO::read (sdpa_io.cpp): skip characters that cannot
// start a number, then read one.
bool next_number(const char*& p, const char* end, double& out) {
  while (p < end && !((*p >= '0' && *p <= '9') || *p == '+' || *p == '-')) ++p;
  if (p >= end) return false;
  char* e = nullptr;
  out = std::strtod(p, &e);
  if (e == p) { ++p; return next_number(p, end, out); }
  p = e;
  return true;
}

}  // namespace

SDPA::SDPA(int device) : device_(device) { setParameterType(PARAMETER_DEFAULT); }

SDPA::~SDPA() { terminate(); }

void SDPA::terminate() {
  if (ctx_) {
    sdpa_gpu_destroy(ctx_);
    ctx_ = nullptr;
  }
  prepared_ = false;
}

// Newton::setNumThreads (sdpa_newton.cpp:2127-2149)
void SDPA::setNumThreads(int numThreads) {
  if (numThreads == 0) {
    const char* env = std::getenv("OMP_NUM_THREADS");
    numThreads_ = env ? std::atoi(env) : 1;
  } else {
    numThreads_ = numThreads;
  }
  if (display_) std::fprintf(display_, "NumThreads  is set as %d\n", numThreads_);
  if (fpout_) std::fprintf(fpout_, "NumThreads  is set as %d\n", numThreads_);
}

// Parameter::setDefaultParameter (sdpa_parts.cpp:182-224)
void SDPA::setParameterType(ParameterType type) {
  // Parameter::setDefaultParameter (sdpa_parts.cpp:192-229)
  paramType_ = type;
  xPrint_ = XPrint_ = YPrint_ = "%+8.3e";
  infPrint_ = "%+10.16e";
  if (type == PARAMETER_STABLE_BUT_SLOW)
    param_ = Param{1000, 1.0e-7, 1.0e+4, 2.0, -1.0e+5, 1.0e+5, 0.1, 0.3, 0.8, 1.0e-7};
  else if (type == PARAMETER_UNSTABLE_BUT_FAST)
    param_ = Param{100, 1.0e-7, 1.0e+2, 2.0, -1.0e+5, 1.0e+5, 0.01, 0.02, 0.95, 1.0e-7};
  else
    param_ = Param{100, 1.0e-7, 1.0e+2, 2.0, -1.0e+5, 1.0e+5, 0.1, 0.2, 0.9, 1.0e-7};
}

void SDPA::inputConstraintNumber(int m) {
  m_ = m;
  c_.assign(m, 0.0);
}
void SDPA::inputBlockNumber(int nBlock) {
  nBlock_ = nBlock;
  blockStruct_.assign(nBlock, 0);
  coneType_.assign(nBlock, SDP);
}
void SDPA::inputBlockSize(int l, int size) {
  if (l <= 0 || l > nBlock_) throw std::runtime_error("inputBlockSize: l out of range");
  blockStruct_[l - 1] = size;
}
void SDPA::inputBlockType(int l, ConeType t) {
  if (l <= 0 || l > nBlock_) throw std::runtime_error("inputBlockType: l out of range");
  if (t == SOCP) throw std::runtime_error("SOCP blocks are not supported (as in SDPA 7.4.4)");
  coneType_[l - 1] = t;
}
void SDPA::inputCVec(int k, double value) {
  if (k <= 0 || k > m_) throw std::runtime_error("inputCVec: k out of range");
  c_[k - 1] = value;
}
void SDPA::initializeUpperTriangleSpace() {
  rk_.clear(); rl_.clear(); ri_.clear(); rj_.clear(); rv_.clear();
}

// SDPA::inputElement (sdpa_call.cpp:285-341): i > j is swapped to the upper triangle
void SDPA::inputElement(int k, int l, int i, int j, double value, bool inputCheck) {
  if (inputCheck) {
    if (k < 0 || k > m_) throw std::runtime_error("inputElement: k out of range");
    if (l <= 0 || l > nBlock_) throw std::runtime_error("inputElement: l out of range");
    int dim = std::abs(blockStruct_[l - 1]);
    if (i <= 0 || i > dim || j <= 0 || j > dim) throw std::runtime_error("inputElement: i/j out of range");
    if (coneType_[l - 1] == LP && i != j) throw std::runtime_error("inputElement: LP entry off diagonal");
  }
  if (i > j) std::swap(i, j);
  rk_.push_back(k); rl_.push_back(l); ri_.push_back(i); rj_.push_back(j); rv_.push_back(value);
}

// SDPA::initializeUpperTriangle (sdpa_call.cpp:650-666): sort by (l, i, j) per k
void SDPA::initializeUpperTriangle(bool check) {
  const size_t n = rk_.size();
  std::vector<size_t> idx(n);
  for (size_t t = 0; t < n; ++t) idx[t] = t;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    if (rk_[a] != rk_[b]) return rk_[a] < rk_[b];
    if (rl_[a] != rl_[b]) return rl_[a] < rl_[b];
    if (ri_[a] != ri_[b]) return ri_[a] < ri_[b];
    return rj_[a] < rj_[b];
  });
  std::vector<int> k(n), l(n), i(n), j(n);
  std::vector<double> v(n);
  for (size_t t = 0; t < n; ++t) {
    k[t] = rk_[idx[t]]; l[t] = rl_[idx[t]]; i[t] = ri_[idx[t]]; j[t] = rj_[idx[t]]; v[t] = rv_[idx[t]];
    if (check && t > 0 && k[t] == k[t - 1] && l[t] == l[t - 1] && i[t] == i[t - 1] && j[t] == j[t - 1])
      throw std::runtime_error("Twice input to the same index");
  }
  rk_.swap(k); rl_.swap(l); ri_.swap(i); rj_.swap(j); rv_.swap(v);
  // block sizes: LP blocks are negative in the internal block structure
  for (int b = 0; b < nBlock_; ++b)
    if (coneType_[b] == LP && blockStruct_[b] > 0) blockStruct_[b] = -blockStruct_[b];
}

void SDPA::loadRecords(int m, int nBlock, const int* bs, const double* c, long nnz, const int* k,
                       const int* l, const int* i, const int* j, const double* v) {
  inputConstraintNumber(m);
  inputBlockNumber(nBlock);
  for (int b = 0; b < nBlock; ++b) {
    blockStruct_[b] = bs[b];
    coneType_[b] = (bs[b] < 0 || bs[b] == 1) ? LP : SDP;
  }
  c_.assign(c, c + m);
  rk_.assign(k, k + nnz); rl_.assign(l, l + nnz); ri_.assign(i, i + nnz); rj_.assign(j, j + nnz);
  rv_.assign(v, v + nnz);
  prepared_ = false;
}

// IO::read for .dat-s and .dat (sdpa_io.cpp:33-66, 473-606)
void SDPA::readInput(const char* filename, FILE* fpout, SparseType type) {
  FILE* f = std::fopen(filename, "rb");
  if (!f) throw std::runtime_error(std::string("cannot open ") + filename);
  std::string buf;
  char tmp[1 << 16];
  size_t got;
  while ((got = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.append(tmp, got);
  std::fclose(f);
  bool sparse = type == SPARSE || (type == AUTO && std::string(filename).size() > 2 &&
                                   std::string(filename).substr(std::string(filename).size() - 2) == "-s");
  if (fpout) std::fprintf(fpout, "data   is %s  : %s\n", filename, sparse ? "sparse" : "dense");
  const char* p = buf.data();
  const char* end = p + buf.size();
  // comment lines start with '*' or '"'
  while (p < end && (*p == '*' || *p == '"')) {
    while (p < end && *p != '\n') ++p;
    if (p < end) ++p;
  }
  double x;
  const char* lineEnd = p;
  while (lineEnd < end && *lineEnd != '\n') ++lineEnd;
  if (!next_number(p, lineEnd, x)) throw std::runtime_error("cannot read m");
  const int m = (int)x;
  p = lineEnd;
  if (!next_number(p, end, x)) throw std::runtime_error("cannot read nBlock");
  const int nBlock = (int)x;
  std::vector<int> bs(nBlock);
  for (int b = 0; b < nBlock; ++b) {
    if (!next_number(p, end, x)) throw std::runtime_error("cannot read block struct");
    bs[b] = (int)x;
  }
  std::vector<double> c(m);
  for (int t = 0; t < m; ++t)
    if (!next_number(p, end, c[t])) throw std::runtime_error("cannot read c");
  std::vector<int> K, L, I, J;
  std::vector<double> V;
  if (sparse) {
    double r[5];
    for (;;) {
      int got5 = 0;
      for (; got5 < 5; ++got5)
        if (!next_number(p, end, r[got5])) break;
      if (got5 < 5) break;
      K.push_back((int)r[0]); L.push_back((int)r[1]); I.push_back((int)r[2]); J.push_back((int)r[3]);
      V.push_back(r[4]);
    }
  } else {
    for (int k = 0; k <= m; ++k)
      for (int b = 0; b < nBlock; ++b) {
        int s = bs[b];
        if (s >= 2) {
          for (int a = 0; a < s; ++a)
            for (int d = 0; d < s; ++d) {
              if (!next_number(p, end, x)) throw std::runtime_error("dense data ended early");
              if (a <= d && x != 0.0) {
                K.push_back(k); L.push_back(b + 1); I.push_back(a + 1); J.push_back(d + 1); V.push_back(x);
              }
            }
        } else {
          int size = std::abs(s);
          for (int a = 0; a < size; ++a) {
            if (!next_number(p, end, x)) throw std::runtime_error("dense data ended early");
            if (x != 0.0) {
              K.push_back(k); L.push_back(b + 1); I.push_back(a + 1); J.push_back(a + 1); V.push_back(x);
            }
          }
        }
      }
  }
  loadRecords(m, nBlock, bs.data(), c.data(), (long)K.size(), K.data(), L.data(), I.data(),
              J.data(), V.data());
}

// Parameter::readFile (sdpa_parts.cpp:231-263): one value per line, positional
void SDPA::readParameter(const char* filename, FILE* fpout) {
  FILE* f = std::fopen(filename, "r");
  if (!f) throw std::runtime_error(std::string("cannot open ") + filename);
  if (fpout) std::fprintf(fpout, "param  is %s \n", filename);
  double v[10];
  char line[512];
  for (int t = 0; t < 10; ++t) {
    if (!std::fgets(line, sizeof(line), f)) { std::fclose(f); throw std::runtime_error("short parameter file"); }
    v[t] = std::strtod(line, nullptr);
  }
  // lines 11-14: the print formats (first token of each line)
  std::string* fmt[4] = {&xPrint_, &XPrint_, &YPrint_, &infPrint_};
  for (int t = 0; t < 4; ++t) {
    if (!std::fgets(line, sizeof(line), f)) break;
    char tok[256];
    if (std::sscanf(line, "%255s", tok) == 1) *fmt[t] = tok;
  }
  std::fclose(f);
  param_ = Param{(int)v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9]};
}

void SDPA::ensureContext() {
  if (!ctx_) check(sdpa_gpu_create(device_, &ctx_), nullptr, "sdpa_gpu_create");
}

void SDPA::buildLayout() {
  // packed offsets per user block in internal order (SDP blocks first, then LP)
  blockOff_.assign(nBlock_, 0);
  long off = 0;
  for (int b = 0; b < nBlock_; ++b)
    if (blockStruct_[b] >= 2) { blockOff_[b] = off; off += (long)blockStruct_[b] * blockStruct_[b]; }
  long lp = off;
  nDim_ = 0;
  for (int b = 0; b < nBlock_; ++b) {
    if (blockStruct_[b] >= 2) nDim_ += blockStruct_[b];
    else { blockOff_[b] = lp; lp += std::abs(blockStruct_[b]); nDim_ += std::abs(blockStruct_[b]); }
  }
  total_ = lp;
}

void SDPA::inputInitXVec(int k, double value) {
  if (k <= 0 || k > m_) throw std::runtime_error("inputInitXVec: k out of range");
  if (initY_.size() != (size_t)m_) initY_.assign(m_, 0.0);
  initY_[k - 1] = -value;  // user x -> -yVec (sdpa_call.cpp:351)
}

// user X -> internal zMat, user Y -> internal xMat (sdpa_call.cpp:354-434)
void SDPA::inputInitXMat(int l, int i, int j, double value) {
  buildLayout();
  if (initZ_.size() != (size_t)total_) initZ_.assign(total_, 0.0);
  int s = blockStruct_[l - 1];
  if (s >= 2) {
    initZ_[blockOff_[l - 1] + (i - 1) + (long)s * (j - 1)] = value;
    initZ_[blockOff_[l - 1] + (j - 1) + (long)s * (i - 1)] = value;
  } else {
    initZ_[blockOff_[l - 1] + i - 1] = value;
  }
}
void SDPA::inputInitYMat(int l, int i, int j, double value) {
  buildLayout();
  if (initX_.size() != (size_t)total_) initX_.assign(total_, 0.0);
  int s = blockStruct_[l - 1];
  if (s >= 2) {
    initX_[blockOff_[l - 1] + (i - 1) + (long)s * (j - 1)] = value;
    initX_[blockOff_[l - 1] + (j - 1) + (long)s * (i - 1)] = value;
  } else {
    initX_[blockOff_[l - 1] + i - 1] = value;
  }
}

void SDPA::setInternalInitPoint(const double* X, const double* Z, const double* y) {
  buildLayout();
  if (!X) { isInitPoint_ = false; initX_.clear(); initZ_.clear(); initY_.clear(); return; }
  initX_.assign(X, X + total_);
  initZ_.assign(Z, Z + total_);
  initY_.assign(m_, 0.0);
  if (y) initY_.assign(y, y + m_);
  isInitPoint_ = true;
}

// structural part of SDPA::initializeSolve (sdpa_solve.cpp:10-29)
void SDPA::initializeSolve() {
  ensureContext();
  buildLayout();
  check(sdpa_gpu_load_problem(ctx_, m_, nBlock_, blockStruct_.data(), c_.data(), (long)rk_.size(),
                              rk_.data(), rl_.data(), ri_.data(), rj_.data(), rv_.data(), kappa_),
        ctx_, "load_problem");
  long info[12];
  sdpa_gpu_problem_info(ctx_, info, 12);
  if (info[5] != total_) throw std::runtime_error("packed layout mismatch");
  prepared_ = true;
}

double SDPA::getDualityGap() const { return objValPrimal_ - objValDual_; }

double SDPA::getDigits() const {
  double mean = (std::fabs(objValPrimal_) + std::fabs(objValDual_)) / 2.0;
  return -std::log10(std::fabs(getDualityGap() / mean));
}

void SDPA::getPhaseString(char* str) const {
  static const char* names[] = {"noINFO", "pFEAS", "dFEAS", "pdFEAS", "pdINF",
                                "pFEAS_dINF", "pINF_dFEAS", "pdOPT", "pUNBD", "dUNBD"};
  std::strcpy(str, names[phaseUser_]);
}

int SDPA::getBlockSize(int l) const { return std::abs(blockStruct_.at(l - 1)); }
SDPA::ConeType SDPA::getBlockType(int l) const { return (ConeType)coneType_.at(l - 1); }

double* SDPA::getResultXMat(int l) {
  if (l <= 0 || l > nBlock_) throw std::runtime_error("getResultXMat: l out of range");
  return Z_.data() + blockOff_[l - 1];
}
double* SDPA::getResultYMat(int l) {
  if (l <= 0 || l > nBlock_) throw std::runtime_error("getResultYMat: l out of range");
  return X_.data() + blockOff_[l - 1];
}

void SDPA::copyCurrentToInit() {
  // sdpa_call.cpp:1238-1245: the final point becomes the next initial point
  initX_ = X_;
  initZ_ = Z_;
  initY_ = y_;
  isInitPoint_ = true;
}


// ---------------------------------------------------------------------------
// Output in the reference's formats (they are the interface): sdpa_io.cpp
// printHeader 608-620, printOneIteration 622-659, printLastInfo 661-773,
// printSolution 844-870; sdpa_struct.cpp Vector::display 104-124,
// DenseMatrix::display 795-835, DenseLinearSpace::displaySolution
// 1624-1660; Parameter::display sdpa_parts.cpp:266-310.

namespace {

const char* kNoPrint = "NOPRINT";

void print_header(FILE* fp) {
  if (!fp) return;
  std::fprintf(fp, "   mu      thetaP  thetaD  objP      objD      alphaP  alphaD  beta \n");
  std::fflush(fp);
}

// one iteration line; the trace holds internal values, the line shows the
// user's primal/dual (REVERSE_PRIMAL_DUAL, sdpa_io.cpp:640-645)
void print_iteration(FILE* fp, int it, const IterationTrace& t) {
  if (!fp) return;
  std::fprintf(fp, "%2d %4.1e %4.1e %4.1e %+7.2e %+7.2e %4.1e %4.1e %4.2e\n", it, t.mu, t.thetaD,
               t.thetaP, -t.objD, -t.objP, t.alphaD, t.alphaP, t.beta);
}

void print_vector(FILE* fp, const double* v, int n, const char* fmt) {
  if (!fp) return;
  if (std::strcmp(fmt, kNoPrint) == 0) { std::fprintf(fp, "%s\n", kNoPrint); return; }
  std::fprintf(fp, "{");
  for (int j = 0; j < n - 1; ++j) { std::fprintf(fp, fmt, v[j]); std::fprintf(fp, ","); }
  if (n > 0) { std::fprintf(fp, fmt, v[n - 1]); std::fprintf(fp, "}\n"); }
  else std::fprintf(fp, "  }\n");
}

// a dense n x n block, column-major, printed row by row
void print_dense(FILE* fp, const double* a, int n, const char* fmt) {
  std::fprintf(fp, "{");
  for (int i = 0; i < n - 1; ++i) {
    std::fprintf(fp, i == 0 ? " " : "  ");
    std::fprintf(fp, "{");
    for (int j = 0; j < n - 1; ++j) { std::fprintf(fp, fmt, a[i + (long)n * j]); std::fprintf(fp, ","); }
    std::fprintf(fp, fmt, a[i + (long)n * (n - 1)]);
    std::fprintf(fp, " },\n");
  }
  if (n > 1) std::fprintf(fp, "  {");
  for (int j = 0; j < n - 1; ++j) { std::fprintf(fp, fmt, a[(n - 1) + (long)n * j]); std::fprintf(fp, ","); }
  std::fprintf(fp, fmt, a[(n - 1) + (long)n * (n - 1)]);
  std::fprintf(fp, " }");
  std::fprintf(fp, n > 1 ? "   }\n" : "\n");
}

const char* phase_name(int p) {
  static const char* names[] = {"noINFO    ", "pFEAS     ", "dFEAS     ", "pdFEAS    ",
                                "pdINF     ", "pFEAS_dINF", "pINF_dFEAS", "pdOPT     ",
                                "pUNBD     ", "dUNBD     "};
  return (p >= 0 && p < 10) ? names[p] : "phase error";
}

// smallest eigenvalue of a symmetric n x n matrix (column-major, overwritten):
// Householder reduction to tridiagonal form, then bisection on the Sturm
// count (the textbook method; the reference calls LAPACK dsyev here)
double sym_min_eig(std::vector<double>& a, int n) {
  if (n == 1) return a[0];
  std::vector<double> d(n), e(n, 0.0), v(n), w(n);
  for (int k = 0; k < n - 2; ++k) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += a[i + (long)n * k] * a[i + (long)n * k];
    alpha = std::sqrt(alpha);
    const double x0 = a[(k + 1) + (long)n * k];
    if (alpha == 0.0) { e[k] = 0.0; continue; }
    const double sgn = x0 >= 0.0 ? 1.0 : -1.0;
    e[k] = -sgn * alpha;
    // v = x - (-sgn alpha) e1, H = I - 2 v v^T / v^T v
    for (int i = 0; i < n; ++i) v[i] = 0.0;
    for (int i = k + 1; i < n; ++i) v[i] = a[i + (long)n * k];
    v[k + 1] += sgn * alpha;
    double vv = 0.0;
    for (int i = k + 1; i < n; ++i) vv += v[i] * v[i];
    if (vv == 0.0) continue;
    // A <- H A H on the trailing block: p = A v * 2/vv, q = p - (v^T p / vv) v
    for (int i = k + 1; i < n; ++i) {
      double t = 0.0;
      for (int j = k + 1; j < n; ++j) t += a[i + (long)n * j] * v[j];
      w[i] = 2.0 * t / vv;
    }
    double vp = 0.0;
    for (int i = k + 1; i < n; ++i) vp += v[i] * w[i];
    const double c = vp / vv;
    for (int i = k + 1; i < n; ++i) w[i] -= c * v[i];
    for (int j = k + 1; j < n; ++j)
      for (int i = k + 1; i < n; ++i) a[i + (long)n * j] -= v[i] * w[j] + w[i] * v[j];
  }
  for (int i = 0; i < n; ++i) d[i] = a[i + (long)n * i];
  e[n - 2] = a[(n - 1) + (long)n * (n - 2)];
  // Gershgorin bounds, then bisection for the smallest eigenvalue
  double lo = d[0], hi = d[0];
  for (int i = 0; i < n; ++i) {
    double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  auto count_below = [&](double x) {
    int cnt = 0;
    double q = d[0] - x;
    if (q < 0.0) ++cnt;
    for (int i = 1; i < n; ++i) {
      const double qq = (q == 0.0) ? 1e-300 : q;
      q = d[i] - x - e[i - 1] * e[i - 1] / qq;
      if (q < 0.0) ++cnt;
    }
    return cnt;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, std::fabs(lo) + std::fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (count_below(mid) >= 1) hi = mid; else lo = mid;
  }
  return 0.5 * (lo + hi);
}

}  // namespace

void SDPA::printSDPAVersion(FILE* fp) {
  if (fp) std::fprintf(fp, "SDPA 7.4.4 interface, sdpa_amd MI355X (gfx950) Newton-direction backend\n");
}

void SDPA::printParametersWith(FILE* fp, const char* f) {
  if (!fp) return;
  if (std::strcmp(f, kNoPrint) == 0) { std::fprintf(fp, "%s\n", kNoPrint); return; }
  const Param& P = param_;
  std::fprintf(fp, "** Parameters **\n");
  std::fprintf(fp, "maxIteration = %d\n", P.maxIteration);
  const char* names[9] = {"epsilonStar  = ", "lambdaStar   = ", "omegaStar    = ",
                          "lowerBound   = ", "upperBound   = ", "betaStar     = ",
                          "betaBar      = ", "gammaStar    = ", "epsilonDash  = "};
  const double vals[9] = {P.epsilonStar, P.lambdaStar, P.omegaStar, P.lowerBound, P.upperBound,
                          P.betaStar, P.betaBar, P.gammaStar, P.epsilonDash};
  for (int t = 0; t < 9; ++t) {
    std::fprintf(fp, "%s", names[t]);
    std::fprintf(fp, f, vals[t]);
    std::fprintf(fp, "\n");
  }
  std::fprintf(fp, "xPrint       = %s \n", xPrint_.c_str());
  std::fprintf(fp, "XPrint       = %s \n", XPrint_.c_str());
  std::fprintf(fp, "YPrint       = %s \n", YPrint_.c_str());
  std::fprintf(fp, "infPrint     = %s \n", infPrint_.c_str());
}

void SDPA::printParameters(FILE* fp) { printParametersWith(fp, "%+10.16e"); }

// internal blocks M (packed) in user block order, as displaySolution does
void SDPA::printBlocks(FILE* fp, const std::vector<double>& M, const char* fmt) {
  if (!fp) return;
  if (std::strcmp(fmt, kNoPrint) == 0) { std::fprintf(fp, "%s\n", kNoPrint); return; }
  std::fprintf(fp, "{\n");
  for (int b = 0; b < nBlock_; ++b) {
    const int s = blockStruct_[b];
    if (M.size() != (size_t)total_) continue;
    if (s >= 2) print_dense(fp, M.data() + blockOff_[b], s, fmt);
    else print_vector(fp, M.data() + blockOff_[b], std::abs(s), fmt);
  }
  std::fprintf(fp, "}\n");
}

void SDPA::printResultXVec(FILE* fp) { print_vector(fp, xVec_.data(), (int)xVec_.size(), xPrint_.c_str()); }
void SDPA::printResultXMat(FILE* fp) { printBlocks(fp, Z_, XPrint_.c_str()); }
void SDPA::printResultYMat(FILE* fp) { printBlocks(fp, X_, YPrint_.c_str()); }

// ComputeTime::display (sdpa_parts.cpp:78-162): the reference's rows and
// format.  Host rows are wall-clock seconds of this solve; the device rows
// (bMat and its parts, gVec, Cholesky, solve, makedX, the inverses, the
// eigenvalue estimates, updateRes) are HIP-event times, recorded when timing
// is enabled (sdpa_gpu_timing), else 0.  Here F2 owners run through the F1
// products (bF1 covers both), the diagonal terms are part of the F3 and F1
// kernels (bDia 0), X and Z are inverted and their step lengths estimated in
// one launch each (the X+Z rows), and file input is not part of solve().
void SDPA::printComputationTime(FILE* fp) {
  if (!fp) return;
  const ComputeTimes& T = computeTime_;
  const double main = T.mainLoop > 0 ? T.mainLoop : 1.0;
  const double eig = T.eigXZ;
  const struct { const char* name; double v; } rows[] = {
      {" Predictor time  =       ", T.predictor},
      {" Corrector time  =       ", T.corrector},
      {" Make bMat time  =       ", T.makebMat},
      {" Make bDia time  =       ", 0.0},
      {" Make bF1  time  =       ", T.bF1},
      {" Make bF2  time  =       ", 0.0},
      {" Make bF3  time  =       ", T.bF3},
      {" Make bPRE time  =       ", 0.0},
      {" Make rMat time  =       ", 0.0},
      {" Make gVec Mul   =       ", 0.0},
      {" Make gVec time  =       ", T.makegVec},
      {" Cholesky bMat   =       ", T.cholesky},
      {" Ste Pre time    =       ", T.stepPredictor},
      {" Ste Cor time    =       ", T.stepCorrector},
      {" solve           =       ", T.solve},
      {" sumDz           =       ", 0.0},
      {" makedX          =       ", T.makedX},
      {" symmetriseDx    =       ", 0.0},
      {" makedXdZ        =       ", 0.0},
      {" xMatTime        =       ", 0.0},
      {" zMatTime        =       ", 0.0},
      {" invzMatTime     =       ", 0.0},
      {" xMatzMatTime    =       ", T.invXZ},
      {" EigxMatTime     =       ", 0.0},
      {" EigzMatTime     =       ", 0.0},
      {" EigxMatzMatTime =       ", eig},
      {" updateRes       =       ", T.updateRes},
      {" EigTime         =       ", eig},
      {" sub_total_bMat  =       ", T.mainLoop - T.makebMat},
      {" Main Loop       =       ", T.mainLoop},
      {" File Check      =       ", 0.0},
      {" File Change     =       ", 0.0},
      {" File Read       =       ", 0.0},
      {" Total           =       ", T.mainLoop},
  };
  std::fprintf(fp, "\n");
  std::fprintf(fp, "                         Time(sec)  Ratio(%% : MainLoop) \n");
  for (const auto& r : rows) std::fprintf(fp, "%s%f,  %f\n", r.name, r.v, r.v / main * 100.0);
  std::fprintf(fp, "\n");
}

// IO::computeDimacs (sdpa_io.cpp:775-813).  The reference's "getOneNorm" is
// the largest absolute entry (sdpa_linear.cpp:184-235); getTwoNorm is the
// squared Euclidean / Frobenius norm (237-268).
void SDPA::getDimacsError(double* err) {
  for (int t = 0; t <= 6; ++t) err[t] = 0.0;
  if (!ctx_ || X_.size() != (size_t)total_) return;
  double b1 = 0.0, c1 = 0.0;
  for (double v : c_) b1 = std::max(b1, std::fabs(v));
  for (size_t e = 0; e < rk_.size(); ++e)
    if (rk_[e] == 0) c1 = std::max(c1, std::fabs(rv_[e]));
  double np, nd;
  check(sdpa_gpu_residuals(ctx_, &np, &nd), ctx_, "residuals");
  std::vector<double> pr(m_), dr(total_);
  check(sdpa_gpu_get_buffer(ctx_, 11, pr.data(), m_), ctx_, "get_buffer");
  check(sdpa_gpu_get_buffer(ctx_, 12, dr.data(), total_), ctx_, "get_buffer");
  double p2 = 0.0, d2 = 0.0, xz = 0.0;
  for (double v : pr) p2 += v * v;
  for (double v : dr) d2 += v * v;
  for (long t = 0; t < total_; ++t) xz += X_[t] * Z_[t];
  auto min_eig = [&](const std::vector<double>& M) {
    double mn = 1.0e50;
    for (int b = 0; b < nBlock_; ++b) {
      const int s = blockStruct_[b];
      if (s >= 2) {
        std::vector<double> a(M.begin() + blockOff_[b], M.begin() + blockOff_[b] + (long)s * s);
        mn = std::min(mn, sym_min_eig(a, s));
      } else {
        for (int t = 0; t < std::abs(s); ++t) mn = std::min(mn, M[blockOff_[b] + t]);
      }
    }
    return mn;
  };
  const double xmin = min_eig(X_), zmin = min_eig(Z_);
  const double ctx = objValPrimal_, bty = objValDual_;
  err[1] = std::sqrt(p2) / (1.0 + b1);
  err[2] = std::max(0.0, -xmin / (1.0 + b1));
  err[3] = std::sqrt(d2) / (1.0 + c1);
  err[4] = std::max(0.0, -zmin / (1.0 + c1));
  err[5] = (ctx - bty) / (1.0 + std::fabs(ctx) + std::fabs(bty));
  err[6] = xz / (1.0 + std::fabs(ctx) + std::fabs(bty));
}

// IO::printDimacs (sdpa_io.cpp:815-842)
void SDPA::printDimacsError(double* e, const char* f, FILE* fp) {
  if (!fp) return;
  const char* tail[6] = {"  [||Ax-b|| / (1+||b||_1)]\n", "  [max(0, -lambda(x)/(1+||b||_1))]\n",
                         "  [||A^Ty + z - c || / (1+||c||_1)]\n", "  [max(0, -lambda(z)/(1+||c||_1))]\n",
                         "  [(<c,x> - <b,y>) / (1 + |<c,x>| + |<b,y>|)]\n",
                         "  [<x,z> / (1 + |<c,x>| + |<b,y>|)]\n"};
  std::fprintf(fp, "\n* DIMACS_ERRORS * \n");
  for (int t = 1; t <= 6; ++t) {
    std::fprintf(fp, "err%d = ", t);
    std::fprintf(fp, f, e[t]);
    std::fprintf(fp, "%s", tail[t - 1]);
  }
}

// SDPA::readInit + IO::read(xMat, yVec, zMat) (sdpa_call.cpp:946-990,
// sdpa_io.cpp:75-228): first the user's x (internal y = -x), then the
// user's X (target 1, internal zMat) and Y (target 2, internal xMat)
void SDPA::readInit(const char* filename, FILE* fpout, SparseType type) {
  FILE* f = std::fopen(filename, "rb");
  if (!f) throw std::runtime_error(std::string("cannot open init file ") + filename);
  std::string buf;
  char tmp[1 << 16];
  size_t got;
  while ((got = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.append(tmp, got);
  std::fclose(f);
  const std::string fn(filename);
  const bool sparse = type == SPARSE || (type == AUTO && fn.size() > 2 && fn.substr(fn.size() - 2) == "-s");
  if (fpout) std::fprintf(fpout, "initial is %s  : %s\n", filename, sparse ? "sparse" : "dense");
  buildLayout();
  initX_.assign(total_, 0.0);
  initZ_.assign(total_, 0.0);
  initY_.assign(m_, 0.0);
  const char* p = buf.data();
  const char* end = p + buf.size();
  double x;
  for (int k = 1; k <= m_; ++k) {
    if (!next_number(p, end, x)) throw std::runtime_error("init file: x vector ended early");
    inputInitXVec(k, x);
  }
  if (sparse) {
    double r[5];
    for (;;) {
      int n5 = 0;
      for (; n5 < 5; ++n5)
        if (!next_number(p, end, r[n5])) break;
      if (n5 < 5) break;
      if (r[4] == 0.0) continue;
      const int target = (int)r[0], l = (int)r[1], i = (int)r[2], j = (int)r[3];
      if (l < 1 || l > nBlock_) throw std::runtime_error("init file: block out of range");
      if (blockStruct_[l - 1] < 2 && i != j)
        throw std::runtime_error("init file: LP entries need row == column");
      if (target == 1) inputInitXMat(l, i, j, r[4]);
      else inputInitYMat(l, i, j, r[4]);
    }
  } else {
    for (int target = 1; target <= 2; ++target)
      for (int b = 0; b < nBlock_; ++b) {
        const int s = blockStruct_[b];
        const int size = s >= 2 ? s * s : std::abs(s);
        for (int e = 0; e < size; ++e) {
          if (!next_number(p, end, x)) throw std::runtime_error("init file: ended early");
          const int i = s >= 2 ? e / s : e, j = s >= 2 ? e % s : e;
          if (x == 0.0 || (s >= 2 && i > j)) continue;
          if (target == 1) inputInitXMat(b + 1, i + 1, j + 1, x);
          else inputInitYMat(b + 1, i + 1, j + 1, x);
        }
      }
  }
}

// SDPA::writeInputSparse (sdpa_call.cpp:1099-1147): the problem as .dat-s
void SDPA::writeInputSparse(const char* filename, const char* fmt) {
  FILE* f = std::fopen(filename, "w");
  if (!f) throw std::runtime_error(std::string("cannot open ") + filename + " to write");
  std::fprintf(f, "%d\n%d\n", m_, nBlock_);
  for (int b = 0; b < nBlock_; ++b) std::fprintf(f, b + 1 < nBlock_ ? "%d," : "%d\n", blockStruct_[b]);
  if (std::strcmp(fmt, kNoPrint) == 0) {
    std::fprintf(f, "%s\n", kNoPrint);
  } else {
    for (int k = 0; k < m_; ++k) { std::fprintf(f, fmt, c_[k]); std::fprintf(f, " "); }
    std::fprintf(f, "\n");
    // F0 first, then F1..Fm; within one matrix the records keep their order
    std::vector<size_t> order(rk_.size());
    for (size_t e = 0; e < order.size(); ++e) order[e] = e;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return rk_[a] < rk_[b]; });
    for (size_t e : order) {
      if (rv_[e] == 0.0) continue;
      std::fprintf(f, "%d %d %d %d ", rk_[e], rl_[e], ri_[e], rj_[e]);
      std::fprintf(f, fmt, rv_[e]);
      std::fprintf(f, "\n");
    }
  }
  std::fclose(f);
}

// SDPA::writeInitSparse (sdpa_call.cpp:1199-1222): the final point as an
// .ini-s (the reference writes minus its final yVec, i.e. the internal y)
void SDPA::writeInitSparse(const char* filename, const char* fmt) {
  FILE* f = std::fopen(filename, "w");
  if (!f) throw std::runtime_error(std::string("cannot open ") + filename + " to write");
  if (std::strcmp(fmt, kNoPrint) == 0) { std::fprintf(f, "%s\n", kNoPrint); std::fclose(f); return; }
  for (int k = 0; k < m_; ++k) { std::fprintf(f, fmt, k < (int)y_.size() ? y_[k] : 0.0); std::fprintf(f, " "); }
  std::fprintf(f, "\n");
  const std::vector<double>* mats[2] = {&Z_, &X_};
  for (int target = 1; target <= 2; ++target) {
    const std::vector<double>& M = *mats[target - 1];
    if (M.size() != (size_t)total_) continue;
    for (int b = 0; b < nBlock_; ++b) {
      const int s = blockStruct_[b];
      if (s < 2) continue;
      for (int i = 0; i < s; ++i)
        for (int j = i; j < s; ++j) {
          const double v = M[blockOff_[b] + i + (long)s * j];
          if (v == 0.0) continue;
          std::fprintf(f, "%d %d %d %d ", target, b + 1, i + 1, j + 1);
          std::fprintf(f, fmt, v);
          std::fprintf(f, "\n");
        }
    }
    for (int b = 0; b < nBlock_; ++b) {
      const int s = blockStruct_[b];
      if (s >= 2) continue;
      for (int i = 0; i < std::abs(s); ++i) {
        std::fprintf(f, "%d %d %d %d ", target, b + 1, i + 1, i + 1);
        std::fprintf(f, fmt, M[blockOff_[b] + i]);
        std::fprintf(f, "\n");
      }
    }
  }
  std::fclose(f);
}

// IO::printLastInfo's summary block for one stream
void SDPA::printLastInfo(FILE* fp) {
  if (!fp) return;
  const double mean = (std::fabs(objValPrimal_) + std::fabs(objValDual_)) / 2.0;
  const double pdgap = std::fabs(objValPrimal_ - objValDual_);
  const double relgap = mean < 1.0 ? pdgap : pdgap / mean;
  const double gap = objValPrimal_ - objValDual_;
  const double digits = -std::log10(std::fabs(pdgap / mean));
  const char* f = infPrint_.c_str();
  auto line = [&](const char* name, double v) {
    std::fprintf(fp, "%s", name);
    std::fprintf(fp, f, v);
    std::fprintf(fp, "\n");
  };
  std::fprintf(fp, "\n");
  std::fprintf(fp, "phase.value  = %s\n", phase_name(phaseUser_));
  std::fprintf(fp, "   Iteration = %d\n", pIteration_);
  line("          mu = ", mu_);
  line("relative gap = ", relgap);
  line("        gap  = ", gap);
  line("     digits  = ", digits);
  line("objValPrimal = ", -objValDual_);
  line("objValDual   = ", -objValPrimal_);
  line("p.feas.error = ", normDualMat_);
  line("d.feas.error = ", normPrimalVec_);
  std::fprintf(fp, "total time   = %.6f\n", solveTime_);
}

// ---------------------------------------------------------------------------
// SDPA::solve, sdpa_solve.cpp:59-233
void SDPA::solve() {
  if (!prepared_) initializeSolve();
  auto t0 = std::chrono::steady_clock::now();
  sdpa_gpu_ctx* g = ctx_;
  const int m = m_;
  const double nDim = nDim_;
  const Param& P = param_;
  trace_.clear();
  failure_ = 0;
  check(sdpa_gpu_timing(g, timing_ ? 1 : 0, nullptr), g, "timing");

  double muInit, muCur;
  if (!isInitPoint_) {
    check(sdpa_gpu_identity_start(g, P.lambdaStar), g, "identity_start");
    muInit = muCur = P.lambdaStar * P.lambdaStar;  // AverageComplementarity::initialize
  } else {
    if (initX_.size() != (size_t)total_ || initZ_.size() != (size_t)total_)
      throw std::runtime_error("initial point incomplete");
    if (initY_.size() != (size_t)m) initY_.assign(m, 0.0);
    check(sdpa_gpu_set_iterate(g, initX_.data(), initY_.data(), initZ_.data()), g, "set_iterate");
    check(sdpa_gpu_keep_initial_point(g), g, "keep_initial_point");
    int code = D_XZ;
    double xz;
    check(sdpa_gpu_dots(g, 1, &code, &xz), g, "dots");
    muInit = muCur = xz / nDim;
    check(sdpa_gpu_inv_chol(g), g, "inv_chol");  // result ignored, as in sdpa_solve.cpp:63
  }
  double normP, normD;
  check(sdpa_gpu_residuals(g, &normP, &normD), g, "residuals");
  const double initNormP = normP, initNormD = normD;
  double betaVal = P.betaStar;
  double thetaP = (initNormP < P.epsilonDash) ? 0.0 : 1.0;
  double thetaD = (initNormD < P.epsilonDash) ? 0.0 : 1.0;
  double rho = 1.0;
  double etaP = P.omegaStar * nDim * muInit, etaD = etaP;
  double objP, objD;
  {
    int codes[2] = {D_CX, D_BY};
    double o[2];
    check(sdpa_gpu_dots(g, 2, codes, o), g, "dots");
    objP = o[0];
    objD = o[1];
  }
  int phase = noINFO;
  double alphaP = 0.0, alphaD = 0.0;

  // Phase::updateCheck (sdpa_parts.cpp:1150-1242)
  auto update_phase = [&]() -> bool {
    const double NONZERO = 1.0e-6, acc = P.epsilonDash;
    int v = noINFO;
    if (normP <= acc) v = (normD <= acc) ? pdFEAS : pFEAS;
    if (v == noINFO && normD <= acc) v = dFEAS;
    phase = v;
    if (v == pdFEAS) {
      double mean = (std::fabs(objP) + std::fabs(objD)) / 2.0;
      double gap = std::fabs(objP - objD);
      double dom = mean < 1.0 ? 1.0 : mean;
      if (gap / dom <= P.epsilonStar) { phase = pdOPT; return false; }
    }
    if (v == noINFO && rho > 1.0 + NONZERO) { phase = pdINF; return false; }
    if (v == pFEAS) {
      if (objP <= -P.upperBound) { phase = pUNBD; return false; }
      if (rho > 1.0 + NONZERO) { phase = pFEAS_dINF; return false; }
    }
    if (v == dFEAS) {
      if (objD >= -P.lowerBound) { phase = dUNBD; return false; }
      if (rho > 1.0 + NONZERO) { phase = pINF_dFEAS; return false; }
    }
    return true;
  };
  update_phase();

  print_header(fpout_);
  print_header(display_);
  // mixed precision: on until the first stall; a stalled refinement is
  // replaced by an FP64 factorization of the same (intact) Schur matrix
  bool mixedOn = precision_ == 1;
  auto solve_schur = [&]() -> bool {
    int sr = sdpa_gpu_solve_bmat(g, 0);
    check(sr, g, "solve_bmat");
    if (sr == 2) {
      mixedOn = false;
      const int inf2 = sdpa_gpu_factor_bmat(g, 0);
      check(inf2, g, "factor_bmat");
      if (inf2 > 0) { failure_ = 1; return false; }
      sr = sdpa_gpu_solve_bmat(g, 0);
      check(sr, g, "solve_bmat");
    }
    return true;
  };
  int it = 0;
  // host wall-clock buckets of ComputeTime (sdpa_parts.h:48-86)
  using clk = std::chrono::steady_clock;
  auto secs = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double>(b - a).count();
  };
  double hPred = 0.0, hCorr = 0.0, hStepP = 0.0, hStepC = 0.0;
  while (update_phase() && it < P.maxIteration) {
    const clk::time_point tPred = clk::now();
    const bool affine = !(phase == noINFO || phase == pFEAS || phase == dFEAS);
    if (phase == pdFEAS) betaVal = 0.0;
    else betaVal = affine ? 2.0 : P.betaBar;
    const int infeasible = (phase == pFEAS || phase == noINFO) ? 1 : 0;

    // Newton::Mehrotra(PREDICTOR)
    check(sdpa_gpu_make_gvec(g, 0, betaVal * muCur, infeasible), g, "make_gvec");
    const bool mixedNow = mixedOn && muCur > muSwitch_;
    // in mixed precision the F1/F2 products of the assembly run in FP32 too
    check(sdpa_gpu_make_bmat(g, mixedNow ? 1 : 0), g, "make_bmat");
    int info = sdpa_gpu_factor_bmat(g, mixedNow ? 1 : 0);
    check(info, g, "factor_bmat");
    if (info > 0) { failure_ = 1; break; }
    if (mixedNow) {
      long st[6];
      check(sdpa_gpu_mixed_stats(g, st), g, "mixed_stats");
      if (!st[5]) mixedOn = false;  // the mixed factorization fell back to FP64
    }
    if (!solve_schur()) break;
    check(sdpa_gpu_make_direction(g, infeasible), g, "make_direction");

    // StepLength::MehrotraPredictor + DirectionParameter::MehrotraCorrector
    const clk::time_point tStepP = clk::now();
    double q[7];
    {
      int codes[5] = {D_CDX, D_BDY, D_XDZ, D_DXZ, D_DXDZ};
      check(sdpa_gpu_dots(g, 5, codes, q), g, "dots");
    }
    alphaP = alphaD = 0.9;
    if (phase == noINFO || phase == dFEAS) {
      if (alphaP > 1.0) alphaP = 1.0;
    } else if (q[0] > 0.0) {
      if (alphaP > alphaD) alphaP = alphaD;
      if (alphaP > 1.0) alphaP = 1.0;
    }
    if (phase == noINFO || phase == pFEAS) {
      if (alphaD > 1.0) alphaD = 1.0;
    } else if (q[1] < 0.0) {
      if (alphaD > alphaP) alphaD = alphaP;
      if (alphaD > 1.0) alphaD = 1.0;
    }
    {
      double muT = muCur + (alphaP * q[3] + alphaD * q[2] + alphaP * alphaD * q[4]) / nDim;
      double v = muT / muCur;
      if (v < 1.0) v = v * v;
      if (phase == pdFEAS) {
        if (v < P.betaStar) v = P.betaStar;
        if (v > 1.0) v = 1.0;
      } else if (v < P.betaBar) {
        v = P.betaBar;
      }
      betaVal = v;
    }

    // Newton::Mehrotra(CORRECTOR)
    const clk::time_point tCorr = clk::now();
    hStepP += secs(tStepP, tCorr);
    hPred += secs(tPred, tCorr);
    check(sdpa_gpu_make_gvec(g, 1, betaVal * muCur, infeasible), g, "make_gvec");
    if (!solve_schur()) break;
    check(sdpa_gpu_make_direction(g, infeasible), g, "make_direction");

    // StepLength::MehrotraCorrector
    const clk::time_point tStepC = clk::now();
    double lamX, lamZ;
    {
      // the eigenvalue estimates and the inner products of the rules below
      // with one host round trip
      int codes[7] = {D_CDX, D_BDY, D_XDZ, D_DXZ, D_DXDZ, D_BY, D_CX};
      check(sdpa_gpu_min_eig_dots(g, &lamX, &lamZ, 7, codes, q), g, "min_eig_dots");
    }
    const double alphaBD = 100.0;
    alphaP = (-lamX > 1.0 / alphaBD) ? -1.0 / lamX : alphaBD;
    alphaD = (-lamZ > 1.0 / alphaBD) ? -1.0 / lamZ : alphaBD;
    alphaP *= P.gammaStar;
    alphaD *= P.gammaStar;
    if (phase == noINFO || phase == dFEAS) {
      if (alphaP > 1.0) alphaP = 1.0;
    } else if (q[0] > 0.0) {
      if (alphaP > alphaD) alphaP = alphaD;
      if (alphaP > 1.0) alphaP = 1.0;
    }
    if (phase == noINFO || phase == pFEAS) {
      if (alphaD > 1.0) alphaD = 1.0;
    } else if (q[1] < 0.0) {
      if (alphaD > alphaP) alphaD = alphaP;
      if (alphaD > 1.0) alphaD = 1.0;
    }
    if (!affine && (phase == noINFO || phase == pFEAS || phase == dFEAS)) {
      const double xi = 3.0;
      double thetaMax = std::max((1.0 - alphaP) * thetaP, (1.0 - alphaD) * thetaD);
      double muNew = muCur + (alphaP * q[3] + alphaD * q[2] + alphaP * alphaD * q[4]) / nDim;
      while (thetaMax * thetaMax * muInit > xi * muNew) {
        double aMax = 0.95 * std::max(alphaP, alphaD);
        alphaP = std::min(alphaP, aMax);
        alphaD = std::min(alphaD, aMax);
        thetaMax = std::max((1.0 - alphaP) * thetaP, (1.0 - alphaD) * thetaD);
        muNew = muCur + (alphaP * q[3] + alphaD * q[2] + alphaP * alphaD * q[4]) / nDim;
        if (alphaP < 1.0e-6 && alphaD < 1.0e-6) break;
      }
    }
    if (phase == pdFEAS) {
      double incD = q[1] * alphaD, incP = q[0] * alphaP;
      double maxRatio = (q[5] - q[6]) / (incP - incD);
      if (maxRatio > 0.0 && maxRatio < 1.0) { alphaP *= maxRatio; alphaD *= maxRatio; }
    }
    {
      const clk::time_point tEnd = clk::now();
      hStepC += secs(tStepC, tEnd);
      hCorr += secs(tCorr, tEnd);
    }
    trace_.push_back(IterationTrace{muCur, thetaP, thetaD, objP, objD, alphaP, alphaD, betaVal});
    print_iteration(fpout_, it, trace_.back());
    print_iteration(display_, it, trace_.back());

    // Solutions::update (sdpa_dataset.cpp:145-171)
    check(sdpa_gpu_update(g, alphaP, alphaD), g, "update");
    if (alphaP < 1.0e-4 && alphaD < 1.0e-4) { failure_ = 2; break; }
    // computeInverse, the inner products and the residual norms with one
    // host round trip
    double o[5];
    {
      int codes[5] = {D_XZ, D_CX, D_BY, isInitPoint_ ? D_X0Z : D_TRZ,
                      isInitPoint_ ? D_XZ0 : D_TRX};
      const int inv = sdpa_gpu_update_end(g, 5, codes, o, &normP, &normD);
      check(inv, g, "update_end");
      if (inv > 0) { failure_ = 3; break; }
    }
    if (!affine) {
      thetaP = std::fabs((1.0 - alphaP) * thetaP);
      thetaD = std::fabs((1.0 - alphaD) * thetaD);
    }
    const double o2 = o[4];
    muCur = o[0] / nDim;
    thetaP = (initNormP > P.epsilonDash * 1.0e-2) ? normP / initNormP : 0.0;
    thetaD = (initNormD > P.epsilonDash * 1.0e-2) ? normD / initNormD : 0.0;
    // SolveInfo::update (sdpa_parts.cpp:914-1060)
    objP = o[1];
    objD = o[2];
    rho = 0.0;
    const double x0z0 = nDim * muInit, xz = nDim * muCur;
    const double x0z = isInitPoint_ ? o[3] : P.lambdaStar * o[3];
    const double xz0 = isInitPoint_ ? o2 : P.lambdaStar * o2;
    const double acc = P.epsilonDash, omega = P.omegaStar;
    if (normP <= acc && xz0 < etaP) etaP = xz0;
    if (normD <= acc && x0z < etaD) etaD = x0z;
    if (normP > acc && normD <= acc)
      rho = thetaP * x0z / ((thetaP + (1.0 - thetaP) * omega) * etaD + xz);
    if (normP <= acc && normD > acc)
      rho = thetaD * xz0 / ((thetaD + (1.0 - thetaD) * omega) * etaP + xz);
    if (normP > acc && normD > acc)
      rho = (thetaD * xz0 + thetaP * x0z) /
            ((thetaP * thetaD + omega * (thetaP * (1.0 - thetaD) + (1.0 - thetaP) * thetaD)) * x0z0 + xz);
    ++it;
  }
  check(sdpa_gpu_residuals(g, &normP, &normD), g, "residuals");
  trace_.push_back(IterationTrace{muCur, thetaP, thetaD, objP, objD, alphaP, alphaD, betaVal});
  pIteration_ = it;
  phase_ = phase;
  phaseUser_ = reverse_phase(phase);
  objValPrimal_ = objP;
  objValDual_ = objD;
  normPrimalVec_ = normP;
  normDualMat_ = normD;
  mu_ = muCur;
  X_.assign(total_, 0.0);
  Z_.assign(total_, 0.0);
  y_.assign(m, 0.0);
  check(sdpa_gpu_get_iterate(g, X_.data(), y_.data(), Z_.data()), g, "get_iterate");
  xVec_.assign(m, 0.0);
  for (int t = 0; t < m; ++t) xVec_[t] = -y_[t];  // sdpa_solve.cpp:200-203
  double tm[10];
  check(sdpa_gpu_timing(g, -1, tm), g, "timing");
  phaseTimes[0] = tm[0]; phaseTimes[1] = tm[2]; phaseTimes[2] = tm[4]; phaseTimes[3] = tm[6];
  phaseTimes[4] = tm[8];
  solveTime_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  {
    double dt[11] = {0};
    check(sdpa_gpu_timing_detail(g, dt, 11), g, "timing_detail");
    ComputeTimes& T = computeTime_;
    T = ComputeTimes();
    T.predictor = hPred;
    T.corrector = hCorr;
    T.stepPredictor = hStepP;
    T.stepCorrector = hStepC;
    T.mainLoop = solveTime_;
    // device buckets (ms -> s), known when timing is enabled
    T.makebMat = dt[0] * 1e-3;
    T.bF1 = dt[1] * 1e-3;
    T.bF3 = (dt[0] - dt[1]) * 1e-3;
    T.cholesky = dt[2] * 1e-3;
    T.solve = dt[3] * 1e-3;
    T.eigXZ = dt[4] * 1e-3;
    T.invXZ = dt[5] * 1e-3;
    T.makegVec = dt[6] * 1e-3;
    T.makedX = dt[7] * 1e-3;
    T.updateRes = (dt[8] + dt[9]) * 1e-3;
  }
  // IO::printLastInfo + printSolution (sdpa_solve.cpp:207-232)
  print_iteration(fpout_, it, trace_.back());
  print_iteration(display_, it, trace_.back());
  printLastInfo(display_);
  printLastInfo(fpout_);
  if (fpout_) {
    printParametersWith(fpout_, infPrint_.c_str());
    printComputationTime(fpout_);
    std::fprintf(fpout_, "xVec = \n");
    printResultXVec(fpout_);
    std::fprintf(fpout_, "xMat = \n");
    printResultXMat(fpout_);
    std::fprintf(fpout_, "yMat = \n");
    printResultYMat(fpout_);
  }
  if (display_) {
    std::fprintf(display_, "  main loop time = %.6f\n", solveTime_);
    std::fprintf(display_, "      total time = %.6f\n", solveTime_);
  }
  if (fpout_) {
    std::fprintf(fpout_, "    main loop time = %.6f\n", solveTime_);
    std::fprintf(fpout_, "        total time = %.6f\n", solveTime_);
  }
}

}  // namespace sdpa_amd

// ---------------------------------------------------------------------------
// C wrapper
struct sdpa_solver {
  sdpa_amd::SDPA* s = nullptr;
  std::string err;
};

extern "C" {

sdpa_solver* sdpa_solver_new(int device) {
  sdpa_solver* h = new sdpa_solver();
  h->s = new sdpa_amd::SDPA(device);
  return h;
}

void sdpa_solver_free(sdpa_solver* h) {
  if (!h) return;
  delete h->s;
  delete h;
}

const char* sdpa_solver_error(sdpa_solver* h) { return h ? h->err.c_str() : "null solver"; }

#define SOLVER_TRY(...)                  \
  if (!h) return -1;                     \
  try {                                  \
    __VA_ARGS__;                         \
  } catch (const std::exception& e) {    \
    h->err = e.what();                   \
    return -2;                           \
  }                                      \
  return 0;

int sdpa_solver_set_param(sdpa_solver* h, const char* name, double v) {
  SOLVER_TRY({
    std::string n(name);
    auto* s = h->s;
    if (n == "maxIteration") s->setParameterMaxIteration((int)v);
    else if (n == "epsilonStar") s->setParameterEpsilonStar(v);
    else if (n == "lambdaStar") s->setParameterLambdaStar(v);
    else if (n == "omegaStar") s->setParameterOmegaStar(v);
    else if (n == "lowerBound") s->setParameterLowerBound(v);
    else if (n == "upperBound") s->setParameterUpperBound(v);
    else if (n == "betaStar") s->setParameterBetaStar(v);
    else if (n == "betaBar") s->setParameterBetaBar(v);
    else if (n == "gammaStar") s->setParameterGammaStar(v);
    else if (n == "epsilonDash") s->setParameterEpsilonDash(v);
    else if (n == "kappa") s->setKappa(v);
    else if (n == "precision") s->setPrecision((int)v, s->muSwitch());
    else if (n == "muSwitch") s->setPrecision(s->precision(), v);
    else throw std::runtime_error("unknown parameter " + n);
  })
}

int sdpa_solver_load(sdpa_solver* h, int m, int nBlock, const int* bs, const double* c, long nnz,
                     const int* k, const int* l, const int* i, const int* j, const double* v) {
  SOLVER_TRY(h->s->loadRecords(m, nBlock, bs, c, nnz, k, l, i, j, v))
}

int sdpa_solver_read_input(sdpa_solver* h, const char* path) { SOLVER_TRY(h->s->readInput(path)) }

int sdpa_solver_read_parameter(sdpa_solver* h, const char* path) {
  SOLVER_TRY(h->s->readParameter(path))
}

int sdpa_solver_set_init(sdpa_solver* h, const double* X, const double* Z, const double* y) {
  SOLVER_TRY(h->s->setInternalInitPoint(X, Z, y))
}

int sdpa_solver_enable_timing(sdpa_solver* h, int on) { SOLVER_TRY(h->s->enableTiming(on != 0)) }

int sdpa_solver_solve(sdpa_solver* h) {
  SOLVER_TRY({
    h->s->initializeSolve();
    h->s->solve();
  })
}

int sdpa_solver_prepare(sdpa_solver* h) { SOLVER_TRY(h->s->initializeSolve()) }

int sdpa_solver_run(sdpa_solver* h) { SOLVER_TRY(h->s->solve()) }

int sdpa_solver_result(sdpa_solver* h, double* out) {
  SOLVER_TRY({
    auto* s = h->s;
    double a = s->getPrimalObj();
    double b = s->getDualObj();
    double mean = (std::fabs(a) + std::fabs(b)) / 2.0;
    double gap = std::fabs(a - b);
    out[0] = a;
    out[1] = b;
    out[2] = s->getPrimalError();
    out[3] = s->getDualError();
    out[4] = s->getMu();
    out[5] = mean < 1.0 ? gap : gap / mean;
    out[6] = s->getIteration();
    out[7] = s->getPhaseValue();
    out[8] = s->failureCode();
    out[9] = s->getSolveTime();
    out[10] = s->getConstraintNumber();
    out[11] = 0;
  })
}

int sdpa_solver_trace(sdpa_solver* h, double* out, int cap) {
  if (!h) return -1;
  const auto& t = h->s->trace();
  int n = (int)t.size() < cap ? (int)t.size() : cap;
  for (int r = 0; r < n; ++r) {
    const auto& x = t[r];
    double row[8] = {x.mu, x.thetaP, x.thetaD, x.objP, x.objD, x.alphaP, x.alphaD, x.beta};
    std::memcpy(out + 8 * r, row, sizeof(row));
  }
  return n;
}

int sdpa_solver_iterate(sdpa_solver* h, double* X, double* Z, double* y) {
  if (!h || !h->s->context()) return -1;
  return sdpa_gpu_get_iterate(h->s->context(), X, y, Z);
}

int sdpa_solver_phase_times(sdpa_solver* h, double* out) {
  if (!h) return -1;
  for (int t = 0; t < 5; ++t) out[t] = h->s->phaseTimes[t];
  return 0;
}

sdpa_gpu_ctx* sdpa_solver_context(sdpa_solver* h) { return h ? h->s->context() : nullptr; }

int sdpa_solver_read_init(sdpa_solver* h, const char* path) {
  SOLVER_TRY(h->s->readInit(path); h->s->setInitPoint(true))
}
int sdpa_solver_write_input(sdpa_solver* h, const char* path, const char* fmt) {
  SOLVER_TRY(h->s->writeInputSparse(path, fmt))
}
int sdpa_solver_write_init(sdpa_solver* h, const char* path, const char* fmt) {
  SOLVER_TRY(h->s->writeInitSparse(path, fmt))
}
int sdpa_solver_dimacs(sdpa_solver* h, double* out) { SOLVER_TRY(h->s->getDimacsError(out)) }
int sdpa_solver_print_result(sdpa_solver* h, int which, const char* path) {
  SOLVER_TRY({
    FILE* f = std::fopen(path, "w");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    if (which == 0) h->s->printResultXVec(f);
    else if (which == 1) h->s->printResultXMat(f);
    else if (which == 2) h->s->printResultYMat(f);
    else h->s->printComputationTime(f);
    std::fclose(f);
  })
}

}  // extern "C"
