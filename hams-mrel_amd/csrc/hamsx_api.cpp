// hamsx_api.cpp -- the C-ABI of include/hamsx.h.
//
// Host side: panel geometry (NormalProcess.f90), wave properties and the
// dispersion roots (ImplementSubs.f90, WavDynSubs.f90), the Newman local
// frames of every source panel, and the orchestration of one frequency
// solve on a context's HIP stream.  Device side: assemble.hip, lu.hip.
#include "hamsx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <complex>
#include <condition_variable>
#include <chrono>
#include <deque>
#include <mutex>
#include <vector>

#include <rocblas/rocblas.h>

#include "hamsx_internal.hpp"
#include "hmatrix_host.hpp"

namespace hamsx {
bool debug_launches() {
  static const bool on = std::getenv("HAMSX_DEBUG") != nullptr;
  return on;
}
void report_launch(const char* name, hipError_t e) {
  if (e != hipSuccess) std::fprintf(stderr, "hamsx: launch of %s failed: %s\n", name, hipGetErrorString(e));
}
}  // namespace hamsx

namespace {
thread_local std::string g_last_error;
}  // namespace

int hamsx::set_error(int code, const char* msg) {
  g_last_error = msg;
  return code;
}

namespace {

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HX_TRY(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return fail(HAMSX_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

constexpr double kPi = 3.141592653589793;   // Const_mod
constexpr double kG = 9.80665;              // Const_mod
constexpr double kGDisp = 9.807;            // DISPERSION's own gravity (WavDynSubs.f90:58)
constexpr int kFrameLen = 24;

double ipow(double a, int b) {  // REAL**INTEGER in flang's square-and-multiply order
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b >>= 1;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

// DISPERSION (WavDynSubs.f90:40-127): first root k tanh(kh) = w^2/g by a
// Newton-Halley iteration on y = kh, then the evanescent roots.
void dispersion(double w, double h, int nrt, double* wvn) {
  const double x = w * w * h / kGDisp;
  double y, t;
  if (x > 0.0 && x <= 2.0) {
    y = std::sqrt(x) * (0.9994 + 0.1701 * x + (double)0.0305f * x * x);
  } else {
    t = x * std::exp(-2.0 * x);
    y = x + 2.0 * t - 6.0 * t * t;
  }
  if (std::fabs(y - x) < (double)1.0e-10f) {
    wvn[0] = x / h;
  } else {
    double exx = 1.0;
    while (exx > 1.0e-10) {
      double trial = y, dnm = trial * trial - x * x;
      double fun = std::log((trial + x) / (trial - x)) / 2.0 - trial;
      double dfun = -x / dnm - 1.0, d2fun = 2.0 * x * trial / (dnm * dnm);
      y = trial - fun / dfun * (1.0 + (fun / dfun) * (d2fun / dfun) / 2.0);
      exx = std::fabs(y - trial);
    }
    wvn[0] = y / h;
  }
  double u = 3.0 * x / (7.0 + 3.0 * x);
  t = 0.0159 + 0.1032 * u + 4.3152 * u * u - 2.8768 * u * u * u;
  for (int i = 2; i <= nrt; i++) {
    double exx = 1.0;
    while (exx > 1.0e-10) {
      double trial = t;
      y = (i - 1) * kPi - trial;
      double dnm = y * y + x * x, fun = std::atan2(x, y) - trial;
      double dfun = x / dnm - 1.0, d2fun = 2.0 * x * trial / (dnm * dnm);
      t = trial - fun / dfun * (1.0 + (fun / dfun) * (d2fun / dfun) / 2.0);
      exx = std::fabs(t - trial);
    }
    y = (i - 1) * kPi - t;
    wvn[i - 1] = y / h;
    t = t - kPi * x / (x * x + kPi * i * (kPi * (i - 1) - t));
  }
}

double norm3(const double* v) { return std::sqrt(ipow(v[0], 2) + ipow(v[1], 2) + ipow(v[2], 2)); }

// Newman local frame of one panel (SingularIntgrMulti.f90:130-181 triangle,
// :326-400 quadrilateral): origin, unit vectors, projected vertices.
void newman_frame(int ncn, const double* xv /* ncn x 3 */, double* fr) {
  std::fill(fr, fr + kFrameLen, 0.0);
  double *po = fr, *ui = fr + 3, *uj = fr + 6, *uk = fr + 9, *ksi = fr + 12, *eta = fr + 17;
  if (ncn == 3) {
    double p12[3], p13[3], po3[3], po1[3], po2[3];
    for (int d = 0; d < 3; d++) {
      p12[d] = xv[3 + d] - xv[d];
      p13[d] = xv[6 + d] - xv[d];
    }
    double n12 = ipow(p12[0], 2) + ipow(p12[1], 2) + ipow(p12[2], 2);
    double alfa = (p12[0] * p13[0] + p12[1] * p13[1] + p12[2] * p13[2]) / n12;
    for (int d = 0; d < 3; d++) po[d] = xv[d] + alfa * p12[d];
    for (int d = 0; d < 3; d++) po3[d] = xv[6 + d] - po[d];
    double l12 = std::sqrt(n12), l3 = norm3(po3);
    for (int d = 0; d < 3; d++) {
      ui[d] = p12[d] / l12;
      uj[d] = po3[d] / l3;
    }
    uk[0] = ui[1] * uj[2] - ui[2] * uj[1];
    uk[1] = ui[2] * uj[0] - ui[0] * uj[2];
    uk[2] = ui[0] * uj[1] - ui[1] * uj[0];
    for (int d = 0; d < 3; d++) {
      po1[d] = xv[d] - po[d];
      po2[d] = xv[3 + d] - po[d];
    }
    ksi[0] = -std::copysign(1.0, alfa) * norm3(po1);
    eta[0] = 0.0;
    ksi[1] = 0.0;
    eta[1] = norm3(po3);
    ksi[2] = norm3(po2);
    eta[2] = 0.0;
    ksi[3] = ksi[0];
    eta[3] = eta[0];
    return;
  }
  double pov[4][3], pon[4][3], un[3];
  for (int d = 0; d < 3; d++) po[d] = (xv[d] + xv[3 + d] + xv[6 + d] + xv[9 + d]) / 4.0;
  for (int c = 0; c < 4; c++)
    for (int d = 0; d < 3; d++) pov[c][d] = xv[3 * c + d] - po[d];
  un[0] = pov[0][1] * pov[1][2] - pov[0][2] * pov[1][1];
  un[1] = pov[0][2] * pov[1][0] - pov[0][0] * pov[1][2];
  un[2] = pov[0][0] * pov[1][1] - pov[0][1] * pov[1][0];
  double li = norm3(pov[0]), lk = norm3(un);
  for (int d = 0; d < 3; d++) {
    ui[d] = pov[0][d] / li;
    uk[d] = un[d] / lk;
  }
  uj[0] = uk[1] * ui[2] - uk[2] * ui[1];
  uj[1] = uk[2] * ui[0] - uk[0] * ui[2];
  uj[2] = uk[0] * ui[1] - uk[1] * ui[0];
  for (int c = 0; c < 4; c++) {
    double dn = uk[0] * pov[c][0] + uk[1] * pov[c][1] + uk[2] * pov[c][2];
    for (int d = 0; d < 3; d++) pon[c][d] = pov[c][d] - dn * uk[d];
  }
  for (int c = 0; c < 4; c++) {
    ksi[c] = ui[0] * pon[c][0] + ui[1] * pon[c][1] + ui[2] * pon[c][2];
    eta[c] = uj[0] * pon[c][0] + uj[1] * pon[c][1] + uj[2] * pon[c][2];
  }
  ksi[4] = ksi[0];
  eta[4] = eta[0];
}

// Allocation epoch of the context whose API call is running on this thread
// (set by set_device): every device (re)allocation bumps it, so a recorded
// launch sequence can tell that a buffer it names has moved.
thread_local unsigned long long* tl_alloc_epoch = nullptr;
inline void bump_epoch() {
  if (tl_alloc_epoch) ++*tl_alloc_epoch;
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  int ensure(size_t count) {
    if (count <= n && p) return 0;
    bump_epoch();
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) return 0;
    if (hipMalloc(&p, sizeof(T) * count) != hipSuccess) {
      p = nullptr;
      return fail(HAMSX_ENOMEM, "hipMalloc of %zu bytes failed", sizeof(T) * count);
    }
    n = count;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Finite-depth pair bins: the pairs (field row i >= row0, source j) of one
// mesh view counting-sorted by the keys of hamsx_internal.hpp, valid for
// (n, nf, h, key); key names the field rows (0: the system's rows).
struct PairBins {
  DevBuf<unsigned> plist;
  DevBuf<int> cnt;
  int n = 0, nf = 0;
  double h = 0.0;
  unsigned long long key = 0;
  unsigned long long gen = 0;  // bumped at every rebuild
  // list slices per region: single entries [reg_begin, reg_end), hull-pair
  // entries [pair_begin, pair_end)
  size_t reg_begin[5] = {0}, reg_end[5] = {0}, pair_begin[5] = {0}, pair_end[5] = {0};
};

}  // namespace

// pinned staging and the device root buffer start with the kNK dispersion
// roots and the four frequency scalars (v, wk, w1, h)
constexpr size_t kStageHead = hamsx::kNK + 4;
// Arnoldi steps the GMRES host loop runs ahead of the device's convergence
// counters (it waits for step j - kGmLag before enqueuing step j + 1)
constexpr int kGmLag = 2;

struct hamsx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int prec = HAMSX_PREC_FP64;
  int refine_iters = 4;       // most correction steps of the mixed path
  double refine_tol = 1.0e-10; // stop when max_q |R_q|_max / |X_q|_max <= this
  // linear solver: HAMSX_SOLVER_LU or HAMSX_SOLVER_GMRES (restart, tolerance, iteration cap)
  int solver = HAMSX_SOLVER_LU;
  int gm_restart = 25, gm_maxiter = 100;
  double gm_tol = 1.0e-4;
  // H-matrix operator of the GMRES path (HMatrix.f90:56-63: leaf 32, eta 2,
  // ACA tolerance 1e-4, rank <= 50); the partition is kept per mesh
  bool hm_on = true;
  double hm_eta = 2.0, hm_tol = 1.0e-4;
  int hm_leaf = 32, hm_rank = 50;
  bool hm_part = false, hm_built = false;
  std::vector<double> hcen;  // hull centroids, n x 3
  std::vector<int> hm_perm;
  std::vector<hamsx::HNode> hm_tree;
  std::vector<hamsx::HBlk> hm_blocks;
  std::vector<int> hm_aca_idx;  // partition index of each ACA block
  DevBuf<int> hmPerm, hmRank, hmLptr, hmLitem;
  DevBuf<int2> hmVxItem, hmVxRed;
  DevBuf<size_t> hmVxOut, hmVxRbase;
  int hm_vx_nred = 0;
  int hm_vx_start[5] = {0, 0, 0, 0, 0};
  DevBuf<int4> hmAdesc, hmDdesc;
  DevBuf<int2> hmLeaves, hmLeafRed;
  DevBuf<int4> hmSeg;
  DevBuf<size_t> hmSegOut, hmLeafRbase;
  DevBuf<double2> hmYpart;
  std::vector<int2> hm_leaves_h;  // the leaves' row ranges, cluster order
  int hm_nseg = 0, hm_leaf_nred = 0;
  DevBuf<size_t> hmUoff, hmVoff, hmRoff, hmToff, hmDoff;
  DevBuf<unsigned char> hmUsed;
  DevBuf<double2> hmU, hmVH, hmD, hmT, hmXp, hmYp;
  DevBuf<float2> hmU32, hmVH32, hmD32;  // mixed precision: the blocks in complex64 for the Arnoldi products
  bool hm_c64 = false;
  int hm_nlr = 0, hm_ndense = 0, hm_nleaf = 0, hm_nrhs = 0;
  double hm_mean_rank = 0.0, hm_ratio = 0.0;
  // mesh; with a lid (irregular-frequency removal) the nf = n + nlid field
  // points are the hull centroids followed by the lid centroids
  int n = 0, nbody = 0, nchunk = 0, nlid = 0, nf = 0;
  DevBuf<double> fx, fy, fz;
  DevBuf<double2> Cm, Xe;  // with a lid: C = A^T A (n x n) and the extended right-hand sides (nf x nrhs)
  rocblas_handle blas = nullptr;
  std::vector<int> body_start;
  DevBuf<double> cx, cy, cz, area, pnsz, nrm, frame;
  DevBuf<int> ncn, body, body_start_d, chunk_body, chunk_j0, chunk_j1;
  // frequency and waves
  hamsx::FreqDev freq{};
  bool have_freq = false;
  DevBuf<double> wvn, beta;
  double* h_pinned = nullptr;  // staging: wvn (NK) + beta + results
  size_t h_pinned_len = 0;
  int nbeta = 0;
  double amp = 1.0, xw0 = 0.0, xw1 = 0.0, rho = 1000.0;
  int isol = 1;
  bool have_waves = false;
  // matrices and workspace
  DevBuf<double2> A, S, X, Xtmp, B, R, D, dum2, partial, ludiag, mvw;
  DevBuf<double2> trinv, trinv2;  // inverted diagonal blocks of the triangular solves (system 2: trinv2)
  DevBuf<float2> Af, Af2;  // complex64 factors (Af2: symmetry system 2)
  DevBuf<int> ipiv, perm, widx, wcnt, flags;
  // finite-depth pair bins of the system's rows
  PairBins bins;
  // the LU's look-ahead stream and events, and the tuning options
  // (hamsx_set_option); lookahead -1: on unless HAMSX_LU_LOOKAHEAD=0
  hamsx::LuAux lu_aux{};
  bool outer_default = true;  // HAMSX_OPT_LU_OUTER_MIN_N never set: kOuterMinNGroup for groups
  int opt_lookahead = -1;
  // symmetry system 2 is factored on its own stream (sym_stream, with its own
  // scratch) while system 1 is factored on `stream`; created on the first
  // hamsx_set_symmetry
  hipStream_t sym_stream = nullptr;
  hipEvent_t sym_go = nullptr, sym_done = nullptr;
  // a frequency group (hamsx_submit_frequency_group): this context's
  // matrix is ready / the group's factorisation is done
  hipEvent_t grp_ev = nullptr;
  DevBuf<int> widx2, wcnt2;
  DevBuf<double2> ludiag2;
  // field points (CalPotentialMulti): nfp points (SoA); vtx: panel vertices,
  // vtx[(3q + d) n + j]; the field view ffx/ffy/ffz holds the hull centroids
  // and then one chunk of field points; field_gen names the point set
  int nfp = 0;
  unsigned long long field_gen = 0;
  DevBuf<double> fpx, fpy, fpz, vtx, ffx, ffy, ffz, fmodetab;
  PairBins fbins;
  DevBuf<double2> fM, fZ, fZ2, fP, fOut, fInc, fDincp;
  DevBuf<int> fOnb;
  // motion equation: the block-diagonal body matrices MATX, BLNR, BQDR, CRS,
  // KSTF (6nb x 6nb each) for bodymat_nbody bodies; motion_quad: max |BQDR| >= 1e-6
  DevBuf<double> bmat;
  int bodymat_nbody = 0, motion_quad = 0;
  DevBuf<double2> mwork, mdspl;
  DevBuf<int> miters;
  DevBuf<double> modetab;  // per-frequency finite-depth mode tables
  // geometric symmetry, one plane (NSYS = 2): sym_mirror 0 = x = 0 (ISX), 1 = y = 0
  // (ISY), -1 none.  The mirror view's points mview = [fx | fy | fz | gx | gy | gz],
  // n + nf each: rows n .. 2n - 1 are the mirrored hull centroids, FLAG taken at
  // the unmirrored ones, then the mirrored lid points (FLAG at the mirrored point).
  // With a lid, Cm2 / Xe2 are system 2's normal matrix and extended right-hand sides.  T2 / S2: the mirrored entries, then system 2's matrix and S-;
  // W1 / W2 the two systems' generator columns; X2 the mirror half's potentials.
  int sym_mirror = -1;
  DevBuf<double> mview, mmodetab;
  PairBins mbins;
  DevBuf<double2> T2, S2, X2, W1, W2, sdinc, Cm2, Xe2, Xtmp2;
  DevBuf<int> ipiv2, perm2;
  // GMRES: Arnoldi blocks, work blocks, per-body preconditioner factors
  DevBuf<double2> gV, gW, gZ, gU, gB, gP, gPt, gDots, pLU;
  DevBuf<double> gNorm, gScale;
  // GMRES Hessenberg / Givens state on the device, its pinned counters
  // (unconverged columns per restart and active columns per Arnoldi step,
  // iterations, worst residual) and the events the host reads them behind
  DevBuf<double2> gH, gG, gSn;
  DevBuf<double> gCs, gBn, gWorst;
  DevBuf<int> gAct, gDone, gMact, gIters, gCnt;
  int* gm_pin = nullptr;
  size_t gm_pin_len = 0;
  hipEvent_t gm_ev[kGmLag + 1] = {};
  DevBuf<int> pPerm, pIpiv, pSz;
  DevBuf<size_t> pLuOff, pXoff, pInvOff;  // per body: factor, packed right-hand-side, inverse-block offsets
  DevBuf<double2> pInv;                   // inverted 32 x 32 diagonal blocks of the body factors
  int pMax = 0;
  std::vector<size_t> pOff;
  DevBuf<double> out, norms, norms2;  // refinement norms (norms2: symmetry system 2)
  bool have_A = false, have_S = false, have_LU = false, have_X = false;
  bool pending = false;  // a submitted frequency solve not yet collected
  bool have_C = false;   // C = A^T A is current (lid only)
  // C formed in complex64 (mixed precision, GMRES, lid, no symmetry plane):
  // its FP64 copy carries complex64 accuracy, so GMRES takes its residuals
  // from A itself; A32 / C32 the complex64 copy of A and the product
  bool c_fp32 = false;
  DevBuf<float2> A32, C32;
  DevBuf<double2> nY;  // A X, (N + L) x nrhs, for those residuals
  int last_iters = 0;
  double last_resid = 0.0;
  // stage timing (measurement hook)
  bool profiling = false;
  // the recorded launch sequence of hamsx_submit_frequency (graphs option):
  // one per context, for the key it was recorded under
  struct GraphKey {
    double h = 0.0;
    bool limit = false;
    unsigned long long epoch = 0, bins = 0;
    bool operator==(const GraphKey& o) const {
      return h == o.h && limit == o.limit && epoch == o.epoch && bins == o.bins;
    }
  };
  GraphKey graph_key_now() const {
    GraphKey k;
    k.h = freq.h;
    k.limit = freq.v == 0.0 || freq.v == -1.0;
    k.epoch = alloc_epoch;
    k.bins = bins.gen;
    return k;
  }
  bool opt_graphs = false;
  bool opt_far32 = true;  // mixed precision: the EIGEN region's evanescent series in FP32
  bool capturing = false;  // recording: the LU runs its one-stream schedule
  unsigned long long alloc_epoch = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  GraphKey graph_key{}, warm_key{};
  bool warm = false, graph_refused = false;
  bool g_have_A = false, g_have_S = false, g_have_LU = false, g_have_X = false, g_have_C = false;
  unsigned long long graph_launches = 0, graph_records = 0;
  // completion notices of submitted solves (hamsx_set_queue)
  hamsx_queue* queue = nullptr;
  int queue_tag = 0;
  struct Mark {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;     // recorded, not yet read
  std::vector<hipEvent_t> pool;
  double stage_ms[HAMSX_NSTAGE] = {0};
  int stage_cnt[HAMSX_NSTAGE] = {0};

  hamsx::MeshDev mesh() const {
    return hamsx::MeshDev{n,      nf,     fx.p,   fy.p, fz.p, cx.p, cy.p, cz.p, area.p, pnsz.p,
                          nrm.p,  ncn.p,  frame.p, body.p, 0,   fx.p, fy.p, fz.p};
  }
  hamsx::MeshDev mirror_view() const {
    hamsx::MeshDev m = mesh();
    const size_t n2 = (size_t)n + nf;
    m.fx = mview.p;
    m.fy = mview.p + n2;
    m.fz = mview.p + 2 * n2;
    m.gx = mview.p + 3 * n2;
    m.gy = mview.p + 4 * n2;
    m.gz = mview.p + 5 * n2;
    m.nf = n + nf;
    m.row0 = n;
    return m;
  }
  bool sym() const { return sym_mirror >= 0; }
  hamsx::WaveDev waves() const { return hamsx::WaveDev{nbeta, beta.p, amp, xw0, xw1, rho, isol}; }
  int nrhs() const { return 6 * nbody + nbeta; }
  size_t out_len() const { return 2 * (size_t)(6 * nbody) * (6 * nbody) + 2 * (size_t)(6 * nbody) * nbeta; }
};

namespace {

int ensure_pinned(hamsx_ctx* c, size_t len) {
  if (len <= c->h_pinned_len) return 0;
  HX_TRY(hipStreamSynchronize(c->stream));  // no copy may still read the old buffer
  bump_epoch();
  if (c->h_pinned) (void)hipHostFree(c->h_pinned);
  c->h_pinned = nullptr;
  c->h_pinned_len = 0;
  HX_TRY(hipHostMalloc((void**)&c->h_pinned, sizeof(double) * len, hipHostMallocDefault));
  c->h_pinned_len = len;
  return 0;
}

hipEvent_t take_event(hamsx_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one stage with events when profiling is on.
struct StageTimer {
  hamsx_ctx* c;
  int stage;
  hipEvent_t a = nullptr;
  StageTimer(hamsx_ctx* ctx, int st) : c(ctx), stage(st) {  // st < 0: not timed
    if (!c->profiling || st < 0) return;
    a = take_event(c);
    if (a) (void)hipEventRecord(a, c->stream);
  }
  ~StageTimer() {
    if (!a) return;
    hipEvent_t b = take_event(c);
    if (!b) return;
    (void)hipEventRecord(b, c->stream);
    c->marks.push_back({stage, a, b});
  }
};

// Reads the recorded event pairs (the stream must have been synchronised).
void harvest_marks(hamsx_ctx* c) {
  for (auto& m : c->marks) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
      c->stage_ms[m.stage] += ms;
      c->stage_cnt[m.stage] += 1;
    }
    c->pool.push_back(m.a);
    c->pool.push_back(m.b);
  }
  c->marks.clear();
}

int set_device(hamsx_ctx* c) {
  tl_alloc_epoch = &c->alloc_epoch;
  HX_TRY(hipSetDevice(c->device));
  return 0;
}

// Drops the recorded launch sequence (any change of configuration).
void drop_graph(hamsx_ctx* c) {
  if (c->graph_exec) (void)hipStreamSynchronize(c->stream);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->warm = false;
  c->graph_refused = false;
}

int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(HAMSX_EDEVICE, "kernel launch: %s", hipGetErrorString(e));
  return 0;
}

// Diagnostic switch: HAMSX_ASSEMBLE_GRID=1 assembles finite depth with the
// unbinned grid kernel (same numbers, no sort; used to cross-check the bins).
bool force_grid_assembly() {
  const char* e = std::getenv("HAMSX_ASSEMBLE_GRID");
  return e && *e && *e != '0';
}

// Diagnostic switch: HAMSX_LIMES_IEEE=1 computes every quotient of the
// epsilon tables (PIDCOCK, EIGENE) by IEEE division instead of the fast
// reciprocal pass (same numbers; used to check that they are).
bool force_limes_ieee() {
  const char* e = std::getenv("HAMSX_LIMES_IEEE");
  return e && *e && *e != '0';
}

// Counting sort of the pairs of mesh view m by (region, NT, FLAG) for the
// current depth; done once per (mesh, rows, h) and reused by every frequency
// at that depth.
int ensure_bins(hamsx_ctx* c, PairBins& B, const hamsx::MeshDev& m, unsigned long long key) {
  using namespace hamsx;
  const double h = c->freq.h;
  if (B.n == m.n && B.nf == m.nf && B.h == h && B.key == key) return 0;
  B.n = 0;
  int rc;
  if ((rc = B.cnt.ensure(kFinBins))) return rc;
  HX_TRY(hipMemsetAsync(B.cnt.p, 0, sizeof(int) * kFinBins, c->stream));
  launch::bin_count(c->stream, m, h, B.cnt.p);
  if ((rc = check_launch())) return rc;
  std::vector<int> cnt(kFinBins), start(kFinBins);
  HX_TRY(hipMemcpyAsync(cnt.data(), B.cnt.p, sizeof(int) * kFinBins, hipMemcpyDeviceToHost, c->stream));
  HX_TRY(hipStreamSynchronize(c->stream));
  size_t off = 0;
  for (int r = 0; r < 5; r++) B.reg_begin[r] = B.reg_end[r] = B.pair_begin[r] = B.pair_end[r] = 0;
  bool seen[2][5] = {};
  for (int k = 0; k < kFinBins; k++) {
    start[k] = (int)off;
    const size_t padded = ((size_t)cnt[k] + 63) / 64 * 64;
    if (cnt[k] > 0) {
      const int r = fin_key_region(k);
      const bool pr = fin_key_paired(k);
      size_t* b = pr ? B.pair_begin : B.reg_begin;
      size_t* e = pr ? B.pair_end : B.reg_end;
      if (!seen[pr][r]) b[r] = off;
      seen[pr][r] = true;
      e[r] = off + padded;
    }
    off += padded;
  }
  if (off > (size_t)INT32_MAX) return fail(HAMSX_EARG, "too many panel pairs for the binned assembly");
  if ((rc = B.plist.ensure(std::max<size_t>(off, 1)))) return rc;
  HX_TRY(hipMemsetAsync(B.plist.p, 0xFF, sizeof(unsigned) * off, c->stream));
  HX_TRY(hipMemcpyAsync(B.cnt.p, start.data(), sizeof(int) * kFinBins, hipMemcpyHostToDevice, c->stream));
  launch::bin_scatter(c->stream, m, h, B.cnt.p, B.plist.p);
  if ((rc = check_launch())) return rc;
  HX_TRY(hipStreamSynchronize(c->stream));  // start[] is read by the copy above
  B.n = m.n;
  B.nf = m.nf;
  B.h = h;
  B.key = key;
  B.gen++;
  return 0;
}

// The binned assembly of mesh view m into A and S (leading dimension
// m.nf - m.row0), regions in decreasing cost so the cheap ones fill the tail.
void assemble_binned(hamsx_ctx* c, const PairBins& B, const hamsx::MeshDev& m, const hamsx::FreqDev& f, double2* A,
                     double2* S, int* ovf, bool timed) {
  hamsx::launch::clear_flags(c->stream, ovf);
  for (int r : {hamsx::kRegLinton, hamsx::kRegPidcock, hamsx::kRegEigene, hamsx::kRegEigen}) {
    StageTimer tr(c, timed ? 4 + r : -1);
    hamsx::launch::assemble_region(c->stream, m, f, r, false, B.plist.p + B.reg_begin[r], B.reg_end[r] - B.reg_begin[r],
                                   A, S, ovf);
    hamsx::launch::assemble_region(c->stream, m, f, r, true, B.plist.p + B.pair_begin[r],
                                   B.pair_end[r] - B.pair_begin[r], A, S, ovf);
  }
}

int ensure_work(hamsx_ctx* c) {
  const size_t n = c->n, nr = std::max(1, c->nrhs());
  int rc;
  if ((rc = c->X.ensure(n * nr)) || (rc = c->Xtmp.ensure(n * nr)) ||
      (rc = c->trinv.ensure(hamsx::launch::tri_inverse_len((int)n))) ||
      (rc = c->dum2.ensure(n * std::max(1, c->nbeta))) ||
      (rc = c->partial.ensure((size_t)c->nchunk * (6 + c->nbeta) * std::max<size_t>(n, c->nf))) ||
      (rc = c->out.ensure(c->out_len() + 8)))
    return rc;
  if (c->sym() && ((rc = c->X2.ensure(n * nr)) || (rc = c->W1.ensure(n * nr)) || (rc = c->W2.ensure(n * nr)) ||
                   (rc = c->sdinc.ensure(2 * n * std::max(1, c->nbeta)))))
    return rc;
  if (c->prec == HAMSX_PREC_MIXED) {
    if ((rc = c->B.ensure(n * nr)) || (rc = c->R.ensure(n * nr)) || (rc = c->D.ensure(n * nr)) ||
        (rc = c->Af.ensure(n * n)) || (rc = c->norms.ensure(2 * nr)) ||
        (rc = c->mvw.ensure(hamsx::launch::matvec_workspace((int)n, (int)nr))))
      return rc;
    if (c->sym() && ((rc = c->Af2.ensure(n * n)) || (rc = c->norms2.ensure(2 * nr)))) return rc;
  }
  return 0;
}

int blas_handle(hamsx_ctx* c) {
  if (!c->blas && rocblas_create_handle(&c->blas) != rocblas_status_success) {
    c->blas = nullptr;
    return fail(HAMSX_EDEVICE, "rocblas_create_handle failed");
  }
  if (rocblas_set_stream(c->blas, c->stream) != rocblas_status_success)
    return fail(HAMSX_EDEVICE, "rocblas_set_stream failed");
  return 0;
}

int do_rhs(hamsx_ctx* c) {
  if (!c->have_S) return fail(HAMSX_ESTATE, "right-hand sides need hamsx_assemble first");
  if (!c->have_waves) return fail(HAMSX_ESTATE, "right-hand sides need hamsx_set_waves first");
  int rc = ensure_work(c);
  if (rc) return rc;
  StageTimer tm(c, 1);
  if (c->sym()) {
    // the two systems' right-hand sides: X = S+ W1, X2 = S- W2
    hamsx::launch::sym_weights(c->stream, c->mesh(), c->freq, c->waves(), c->nbody, c->sym_mirror, c->sdinc.p,
                               c->W1.p, c->W2.p);
    if ((rc = check_launch()) || (rc = blas_handle(c))) return rc;
    // with a lid the products run over all nf rows and then D = A_ip^T B_ip
    const bool lid = c->nlid > 0;
    if (lid && ((rc = c->Xe.ensure((size_t)c->nf * c->nrhs())) || (rc = c->Xe2.ensure((size_t)c->nf * c->nrhs()))))
      return rc;
    const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
    for (int ip = 0; ip < 2; ip++) {
      double2* dst = lid ? (ip == 0 ? c->Xe.p : c->Xe2.p) : (ip == 0 ? c->X.p : c->X2.p);
      if (rocblas_zgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c->nf, c->nrhs(), c->n, &one,
                        reinterpret_cast<const rocblas_double_complex*>(ip == 0 ? c->S.p : c->S2.p), c->nf,
                        reinterpret_cast<const rocblas_double_complex*>(ip == 0 ? c->W1.p : c->W2.p), c->n, &zero,
                        reinterpret_cast<rocblas_double_complex*>(dst), c->nf) != rocblas_status_success)
        return fail(HAMSX_EDEVICE, "rocblas_zgemm (symmetric right-hand sides) failed");
      if (lid && rocblas_zgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, c->n, c->nrhs(), c->nf,
                               &one, reinterpret_cast<const rocblas_double_complex*>(ip == 0 ? c->A.p : c->T2.p),
                               c->nf, reinterpret_cast<const rocblas_double_complex*>(dst), c->nf, &zero,
                               reinterpret_cast<rocblas_double_complex*>(ip == 0 ? c->X.p : c->X2.p),
                               c->n) != rocblas_status_success)
        return fail(HAMSX_EDEVICE, "rocblas_zgemm (A^T B, symmetric) failed");
    }
    return 0;
  }
  const bool lid = c->nlid > 0;
  if (lid && (rc = c->Xe.ensure((size_t)c->nf * c->nrhs()))) return rc;
  hamsx::launch::rhs(c->stream, c->mesh(), c->freq, c->waves(), c->nbody, c->body_start.data(), c->body_start_d.p,
                     c->S.p, c->dum2.p, c->partial.p, c->nchunk, c->chunk_body.p, c->chunk_j0.p, c->chunk_j1.p,
                     lid ? c->Xe.p : c->X.p);
  if ((rc = check_launch())) return rc;
  if (lid) {  // D = A^T B, plain transpose (ASSB_RBC/DBC_IRR_MULTI, AssbMatx_irrMulti.f90:300-316, :447-461)
    if ((rc = blas_handle(c))) return rc;
    const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
    if (rocblas_zgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, c->n, c->nrhs(), c->nf, &one,
                      reinterpret_cast<const rocblas_double_complex*>(c->A.p), c->nf,
                      reinterpret_cast<const rocblas_double_complex*>(c->Xe.p), c->nf, &zero,
                      reinterpret_cast<rocblas_double_complex*>(c->X.p), c->n) != rocblas_status_success)
      return fail(HAMSX_EDEVICE, "rocblas_zgemm (A^T B) failed");
  }
  return 0;
}

// With a lid: C = A^T A (ASSB_LEFT_IRR_MULTI, AssbMatx_irrMulti.f90:154-171),
// the matrix the LU, the refinement and GMRES then work on.  The plain
// transpose makes C complex symmetric, so only its lower triangle is
// computed (zsyrk, half the flops of the full product) and then mirrored.
// Diagnostic switch: HAMSX_NORMAL_FP64=1 forms C = A^T A in FP64 on the
// mixed-precision GMRES path too (for comparisons).
bool force_normal_fp64() {
  const char* e = std::getenv("HAMSX_NORMAL_FP64");
  return e && *e && *e != '0';
}

// C = A^T A (plain transpose, AssbMatx_irrMulti.f90:154-171) with a lid.  In
// mixed precision on the GMRES path (no symmetry plane) the product runs in
// complex64 on FP32 MFMA (csyrk of a complex64 copy of A) and is widened to
// FP64 for the H-matrix compression and the preconditioner: the Arnoldi
// products read complex64 blocks anyway, and every residual GMRES tests is
// taken from A itself in FP64 (normal_residual), so GMRES still converges
// on the FP64 normal equations.
int ensure_normal_matrix(hamsx_ctx* c) {
  if (c->nlid == 0) return 0;
  const bool want32 = c->prec == HAMSX_PREC_MIXED && c->solver == HAMSX_SOLVER_GMRES && !c->sym() &&
                      !force_normal_fp64();
  if (c->have_C && c->c_fp32 == want32) return 0;
  c->have_C = false;
  int rc;
  if ((rc = c->Cm.ensure((size_t)c->n * c->n)) || (rc = blas_handle(c))) return rc;
  if (want32) {
    const size_t na = (size_t)c->nf * c->n, nc = (size_t)c->n * c->n;
    if ((rc = c->A32.ensure(na)) || (rc = c->C32.ensure(nc))) return rc;
    hamsx::launch::to_c64(c->stream, c->A.p, c->A32.p, na);
    const rocblas_float_complex onef(1.0f, 0.0f), zerof(0.0f, 0.0f);
    if (rocblas_csyrk(c->blas, rocblas_fill_lower, rocblas_operation_transpose, c->n, c->nf, &onef,
                      reinterpret_cast<const rocblas_float_complex*>(c->A32.p), c->nf, &zerof,
                      reinterpret_cast<rocblas_float_complex*>(c->C32.p), c->n) != rocblas_status_success)
      return fail(HAMSX_EDEVICE, "rocblas_csyrk (A^T A in complex64) failed");
    hamsx::launch::to_c128(c->stream, c->C32.p, c->Cm.p, nc);
    hamsx::launch::mirror_lower(c->stream, c->Cm.p, c->n);
    if ((rc = check_launch())) return rc;
    c->c_fp32 = true;
    c->have_C = true;
    return 0;
  }
  c->c_fp32 = false;
  if (c->sym() && (rc = c->Cm2.ensure((size_t)c->n * c->n))) return rc;
  const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
  for (int ip = 0; ip < (c->sym() ? 2 : 1); ip++) {  // with a symmetry plane: C2 = A2^T A2 as well
    double2* C = ip == 0 ? c->Cm.p : c->Cm2.p;
    if (rocblas_zsyrk(c->blas, rocblas_fill_lower, rocblas_operation_transpose, c->n, c->nf, &one,
                      reinterpret_cast<const rocblas_double_complex*>(ip == 0 ? c->A.p : c->T2.p), c->nf, &zero,
                      reinterpret_cast<rocblas_double_complex*>(C), c->n) != rocblas_status_success)
      return fail(HAMSX_EDEVICE, "rocblas_zsyrk (A^T A) failed");
    hamsx::launch::mirror_lower(c->stream, C, c->n);
  }
  if ((rc = check_launch())) return rc;
  c->have_C = true;
  return 0;
}

// The LU's schedule and shape for this context: the look-ahead stream unless
// switched off (HAMSX_OPT_LU_LOOKAHEAD, or HAMSX_LU_LOOKAHEAD=0 by default:
// one stream, the same factors bit for bit) or `allow` is false, and the
// selection-block and outer-block options (for a batch of two or more
// matrices, group, outer blocks from kOuterMinNGroup on unless the option
// was set).
hamsx::LuAux lu_options(hamsx_ctx* c, bool allow = true, bool group = false) {
  hamsx::LuAux a = c->lu_aux;
  if (group && c->outer_default) a.outer_min_n = hamsx::kOuterMinNGroup;
  bool on = c->opt_lookahead != 0;
  if (c->opt_lookahead < 0) {
    const char* e = std::getenv("HAMSX_LU_LOOKAHEAD");
    on = !(e && *e == '0');
  }
  if (!on || !allow || c->capturing) a.s2 = nullptr;
  return a;
}


// Symmetry system 2's stream, or null when HAMSX_SYM_STREAMS=0 (diagnostic:
// both systems on the context stream, one after the other; the same results
// bit for bit).
hipStream_t sym_second_stream(hamsx_ctx* c) {
  const char* e = std::getenv("HAMSX_SYM_STREAMS");
  if (e && *e == '0') return nullptr;
  return c->sym_stream;
}

// the square matrix the solvers work on: A, or C = A^T A with a lid
double2* system_matrix(hamsx_ctx* c) { return c->nlid > 0 ? c->Cm.p : c->A.p; }
// symmetry system 2's: A2 (in T2), or C2 = A2^T A2 with a lid
double2* system_matrix2(hamsx_ctx* c) { return c->nlid > 0 ? c->Cm2.p : c->T2.p; }

// The factorisation's buffers: pivots, permutation, tournament workspace,
// diagonal-block scratch (and the solves' workspace).
int factor_buffers(hamsx_ctx* c) {
  int rc = ensure_work(c);
  if (rc) return rc;
  const int n = c->n;
  const int nsets0 = (n + hamsx::kSelRows - 1) / hamsx::kSelRows;
  if ((rc = c->ipiv.ensure(n)) || (rc = c->perm.ensure(n)) ||
      (rc = c->widx.ensure((size_t)2 * nsets0 * hamsx::kPanelNB)) || (rc = c->wcnt.ensure((size_t)2 * nsets0)) ||
      (rc = c->ludiag.ensure((size_t)hamsx::kPanelNB * hamsx::kPanelNB)))
    return rc;
  if (c->sym() && ((rc = c->ipiv2.ensure(n)) || (rc = c->perm2.ensure(n)) ||
                   (rc = c->widx2.ensure((size_t)2 * nsets0 * hamsx::kPanelNB)) ||
                   (rc = c->wcnt2.ensure((size_t)2 * nsets0)) ||
                   (rc = c->ludiag2.ensure((size_t)hamsx::kPanelNB * hamsx::kPanelNB))))
    return rc;
  return 0;
}

int do_factor(hamsx_ctx* c) {
  if (c->have_LU) return 0;
  if (!c->have_A) return fail(HAMSX_ESTATE, "hamsx_factor needs hamsx_assemble first");
  int rc = factor_buffers(c);
  if (rc) return rc;
  const int n = c->n;
  StageTimer tm(c, 2);
  if (c->sym()) {
    // the two systems A +- T2, or their normal matrices (INFO of system 2 in
    // flags[5]): system 2 on sym_stream with its own scratch, concurrently
    // with system 1 on the context stream, which then waits for it.  Both
    // factorisations are latency-bound chains, so they overlap well; neither
    // takes the look-ahead stream, which keeps two streams busy per context
    // (with four contexts in flight, more streams than hardware queues
    // serialise them: measured slower).
    if ((rc = ensure_normal_matrix(c))) return rc;
    const bool two = sym_second_stream(c) != nullptr;
    hipStream_t s2 = two ? c->sym_stream : c->stream;
    if (two) {
      HX_TRY(hipEventRecord(c->sym_go, c->stream));
      HX_TRY(hipStreamWaitEvent(s2, c->sym_go, 0));
    }
    int* widx2 = two ? c->widx2.p : c->widx.p;
    int* wcnt2 = two ? c->wcnt2.p : c->wcnt.p;
    double2* diag2 = two ? c->ludiag2.p : c->ludiag.p;
    const hamsx::LuAux o1 = lu_options(c, !two);
    const hamsx::LuAux* aux1 = &o1;
    if (c->prec == HAMSX_PREC_FP64) {
      hamsx::launch::lu_factor(s2, system_matrix2(c), n, c->ipiv2.p, c->perm2.p, widx2, wcnt2, c->flags.p + 5,
                               diag2, aux1);
      hamsx::launch::lu_factor(c->stream, system_matrix(c), n, c->ipiv.p, c->perm.p, c->widx.p, c->wcnt.p,
                               c->flags.p + 1, c->ludiag.p, aux1);
      if (c->nlid == 0) c->have_A = false;
    } else {  // complex64 factors; the FP64 matrices stay for the residuals
      hamsx::launch::to_c64(s2, system_matrix2(c), c->Af2.p, (size_t)n * n);
      hamsx::launch::lu_factor_c64(s2, c->Af2.p, n, c->ipiv2.p, c->perm2.p, widx2, wcnt2, c->flags.p + 5, diag2,
                                   aux1);
      hamsx::launch::to_c64(c->stream, system_matrix(c), c->Af.p, (size_t)n * n);
      hamsx::launch::lu_factor_c64(c->stream, c->Af.p, n, c->ipiv.p, c->perm.p, c->widx.p, c->wcnt.p,
                                   c->flags.p + 1, c->ludiag.p, aux1);
    }
    if (two) {
      HX_TRY(hipEventRecord(c->sym_done, s2));
      HX_TRY(hipStreamWaitEvent(c->stream, c->sym_done, 0));
    }
    c->have_LU = true;
    return check_launch();
  }
  if ((rc = ensure_normal_matrix(c))) return rc;
  double2* M = system_matrix(c);
  const hamsx::LuAux o = lu_options(c);
  if (c->prec == HAMSX_PREC_FP64) {
    hamsx::launch::lu_factor(c->stream, M, n, c->ipiv.p, c->perm.p, c->widx.p, c->wcnt.p, c->flags.p + 1,
                             c->ludiag.p, &o);
    if (c->nlid == 0) c->have_A = false;  // factored in place (with a lid, C is factored and A kept)
  } else {
    hamsx::launch::to_c64(c->stream, M, c->Af.p, (size_t)n * n);
    hamsx::launch::lu_factor_c64(c->stream, c->Af.p, n, c->ipiv.p, c->perm.p, c->widx.p, c->wcnt.p,
                                 c->flags.p + 1, c->ludiag.p, &o);
  }
  c->have_LU = true;
  return check_launch();
}

// ---------------------------------------------------------------------------
// Restarted GMRES (ISOLV = 2 path of the reference: ZBLOCK_GMRES_SOLVE /
// ZGMRES_SOLVE, src/IterativeSolvers.f90:324-562, block-diagonal LU
// preconditioner SETUP_BLOCK_PRECONDITIONER :190-307), right-preconditioned,
// x0 = 0, all right-hand sides advancing together; the Hessenberg / Givens
// work runs on the device and the host waits once per restart cycle (and
// reads the active-column count kGmLag steps behind).  The operator is the assembled dense
// matrix (the reference's H-matrix samples that same matrix,
// src/HMatrix.f90:633-665).  Rotations are the standard complex Givens
// rotations with real cosine, which annihilate the subdiagonal for complex
// Hessenberg entries.
// ---------------------------------------------------------------------------
namespace {
int gmres_precond_setup(hamsx_ctx* c, const double2* M) {
  const int nb = c->nbody;
  c->pOff.assign(nb + 1, 0);
  size_t maxb = 0;
  for (int b = 0; b < nb; b++) {
    const size_t sz = (size_t)(c->body_start[b + 1] - c->body_start[b]);
    c->pOff[b + 1] = c->pOff[b] + sz * sz;
    maxb = std::max(maxb, sz);
  }
  int rc;
  const int nsets0 = ((int)maxb + hamsx::kSelRows - 1) / hamsx::kSelRows;
  if ((rc = c->pLU.ensure(c->pOff[nb])) || (rc = c->pPerm.ensure(c->n)) || (rc = c->pIpiv.ensure(maxb)) ||
      (rc = c->widx.ensure((size_t)2 * nsets0 * hamsx::kPanelNB)) || (rc = c->wcnt.ensure((size_t)2 * nsets0)) ||
      (rc = c->ludiag.ensure((size_t)hamsx::kPanelNB * hamsx::kPanelNB)))
    return rc;
  for (int b = 0; b < nb; b++) {
    const int r0 = c->body_start[b], sz = c->body_start[b + 1] - r0;
    double2* blk = c->pLU.p + c->pOff[b];
    hamsx::launch::block_copy(c->stream, M, c->n, r0, sz, blk);
    hamsx::launch::lu_factor(c->stream, blk, sz, c->pIpiv.p, c->pPerm.p + r0, c->widx.p, c->wcnt.p,
                             c->flags.p + 1, c->ludiag.p);
  }
  // descriptors of the batched block solves
  const int nr = c->nrhs();
  std::vector<int> sz(nb);
  std::vector<size_t> xo(nb);
  for (int b = 0; b < nb; b++) {
    sz[b] = c->body_start[b + 1] - c->body_start[b];
    xo[b] = (size_t)nr * c->body_start[b];
  }
  c->pMax = (int)maxb;
  // the inverted 32 x 32 diagonal blocks of every body's factors
  std::vector<size_t> io(nb + 1, 0);
  for (int b = 0; b < nb; b++) io[b + 1] = io[b] + hamsx::launch::tri_inverse_len(sz[b]);
  if ((rc = c->pSz.ensure(nb)) || (rc = c->pLuOff.ensure(nb)) || (rc = c->pXoff.ensure(nb)) ||
      (rc = c->pInvOff.ensure(nb)) || (rc = c->pInv.ensure(std::max<size_t>(io[nb], 1))))
    return rc;
  HX_TRY(hipMemcpyAsync(c->pSz.p, sz.data(), sizeof(int) * nb, hipMemcpyHostToDevice, c->stream));
  HX_TRY(hipMemcpyAsync(c->pLuOff.p, c->pOff.data(), sizeof(size_t) * nb, hipMemcpyHostToDevice, c->stream));
  HX_TRY(hipMemcpyAsync(c->pXoff.p, xo.data(), sizeof(size_t) * nb, hipMemcpyHostToDevice, c->stream));
  HX_TRY(hipMemcpyAsync(c->pInvOff.p, io.data(), sizeof(size_t) * nb, hipMemcpyHostToDevice, c->stream));
  hamsx::launch::tri_inverse_systems(c->stream, nb, c->pMax, c->pLU.p, c->pLuOff.p, c->pSz.p, c->pInvOff.p, c->pInv.p);
  HX_TRY(hipStreamSynchronize(c->stream));  // the host vectors go out of scope
  return check_launch();
}

// Z = M^-1 V (n x nrhs): all bodies' block solves in one batched pass
void gmres_precond_apply(hamsx_ctx* c, const double2* V, double2* Z, int nrhs) {
  hamsx::launch::lu_solve_systems(c->stream, c->nbody, c->pMax, c->pLU.p, c->pLuOff.p, c->pSz.p, c->pInv.p,
                                  c->pInvOff.p, c->pXoff.p, c->body_start_d.p, c->body.p, c->pPerm.p, c->n, V, Z,
                                  c->gP.p, c->gPt.p, nrhs);
}

// ---------------------------------------------------------------------------
// H-matrix operator (HMATRIX_BUILD, HMatrix.f90:135-199): the partition of
// the hull centroids once per mesh and parameters, then per frequency the
// ACA of the admissible blocks of the operator on the device, the choice
// dense / low-rank by the reference's storage test (rank (nr + nc) >= nr nc
// stays dense, :731-737), the dense copies and the per-leaf block lists of
// the product.
//
// ACA's stopping norm: the reference updates ||U VH||_F^2 with
// sum_J Re(conj(U(J,l)) U(J,k)) Re(conj(VH(k,J)) VH(l,J)) over J = 1..NR
// (:703-709), a product of real parts entry by entry that also reads VH past
// its NC columns when NR > NC.  The kernel adds the exact cross terms
// 2 Re(<u_l, u_k> <v_k, v_l>) instead, so its ranks follow the stated
// criterion |u_k| |v_k| < tol ||U VH||_F.
// ---------------------------------------------------------------------------
int hm_build(hamsx_ctx* c, const double2* Aop, int nrhs) {
  using hamsx::HBlk;
  const int n = c->n;
  int rc;
  if (!c->hm_part) {
    hamsx::hmatrix_partition(n, c->hcen.data(), c->hm_leaf, c->hm_eta, c->hm_perm, c->hm_tree, c->hm_blocks);
    if ((rc = c->hmPerm.ensure(n))) return rc;
    HX_TRY(hipMemcpyAsync(c->hmPerm.p, c->hm_perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    c->hm_aca_idx.clear();
    for (size_t k = 0; k < c->hm_blocks.size(); k++)
      if (c->hm_blocks[k].aca) c->hm_aca_idx.push_back((int)k);
    // leaves in cluster order
    std::vector<int2> leaves;
    for (const auto& nd : c->hm_tree)
      if (nd.left < 0) leaves.push_back(make_int2(nd.s, nd.e + 1));
    std::sort(leaves.begin(), leaves.end(), [](const int2& a, const int2& b) { return a.x < b.x; });
    c->hm_nleaf = (int)leaves.size();
    c->hm_leaves_h = leaves;
    if ((rc = c->hmLeaves.ensure(leaves.size()))) return rc;
    HX_TRY(hipMemcpyAsync(c->hmLeaves.p, leaves.data(), sizeof(int2) * leaves.size(), hipMemcpyHostToDevice,
                          c->stream));
    HX_TRY(hipStreamSynchronize(c->stream));
    c->hm_part = true;
  }
  // ACA of the admissible blocks: worst-case storage for rank min(rank_max, nr, nc)
  const int na = (int)c->hm_aca_idx.size();
  std::vector<int4> adesc(na);
  std::vector<size_t> uoff(na), voff(na), roff(na);
  size_t us = 0, vs = 0, rsum = 0;
  for (int a = 0; a < na; a++) {
    const HBlk& b = c->hm_blocks[c->hm_aca_idx[a]];
    const size_t km = (size_t)std::min(c->hm_rank, std::min(b.nr, b.nc));
    adesc[a] = make_int4(b.rs, b.nr, b.cs, b.nc);
    uoff[a] = us;
    voff[a] = vs;
    roff[a] = rsum;
    us += (size_t)b.nr * km;
    vs += km * (size_t)b.nc;
    rsum += (size_t)b.nr;
  }
  if ((rc = c->hmAdesc.ensure(std::max(na, 1))) || (rc = c->hmUoff.ensure(std::max(na, 1))) ||
      (rc = c->hmVoff.ensure(std::max(na, 1))) || (rc = c->hmRoff.ensure(std::max(na, 1))) ||
      (rc = c->hmRank.ensure(std::max(na, 1))) || (rc = c->hmToff.ensure(std::max(na, 1))) ||
      (rc = c->hmUsed.ensure(std::max<size_t>(rsum, 1))) || (rc = c->hmU.ensure(std::max<size_t>(us, 1))) ||
      (rc = c->hmVH.ensure(std::max<size_t>(vs, 1))))
    return rc;
  if (na > 0) {
    HX_TRY(hipMemcpyAsync(c->hmAdesc.p, adesc.data(), sizeof(int4) * na, hipMemcpyHostToDevice, c->stream));
    HX_TRY(hipMemcpyAsync(c->hmUoff.p, uoff.data(), sizeof(size_t) * na, hipMemcpyHostToDevice, c->stream));
    HX_TRY(hipMemcpyAsync(c->hmVoff.p, voff.data(), sizeof(size_t) * na, hipMemcpyHostToDevice, c->stream));
    HX_TRY(hipMemcpyAsync(c->hmRoff.p, roff.data(), sizeof(size_t) * na, hipMemcpyHostToDevice, c->stream));
  }
  hamsx::launch::hm_aca(c->stream, Aop, n, c->hmPerm.p, na, c->hmAdesc.p, c->hmUoff.p, c->hmVoff.p, c->hmRoff.p,
                        c->hmUsed.p, c->hmU.p, c->hmVH.p, c->hmRank.p, c->hm_rank, c->hm_tol);
  if ((rc = check_launch())) return rc;
  std::vector<int> rank(na);
  if (na > 0)
    HX_TRY(hipMemcpyAsync(rank.data(), c->hmRank.p, sizeof(int) * na, hipMemcpyDeviceToHost, c->stream));
  HX_TRY(hipStreamSynchronize(c->stream));
  // storage test, dense list, per-block kind (>= 0 dense index, < 0 ~low-rank index)
  std::vector<int> kind(c->hm_blocks.size());
  std::vector<int4> ddesc;
  std::vector<size_t> doff, toff(na, 0);
  size_t ds = 0, ts = 0, lr_words = 0;
  long long rank_sum = 0;
  int nlr = 0;
  std::vector<int> aca_pos(c->hm_blocks.size(), -1);
  for (int a = 0; a < na; a++) aca_pos[c->hm_aca_idx[a]] = a;
  for (size_t k = 0; k < c->hm_blocks.size(); k++) {
    const HBlk& b = c->hm_blocks[k];
    const int a = aca_pos[k];
    if (a >= 0 && rank[a] > 0 && (long long)rank[a] * (b.nr + b.nc) < (long long)b.nr * b.nc) {
      kind[k] = ~a;
      toff[a] = ts;
      ts += (size_t)rank[a] * nrhs;
      lr_words += (size_t)rank[a] * (b.nr + b.nc);
      rank_sum += rank[a];
      nlr++;
      continue;
    }
    if (a >= 0) rank[a] = 0;  // kept dense: no product through its factors
    kind[k] = (int)ddesc.size();
    ddesc.push_back(make_int4(b.rs, b.nr, b.cs, b.nc));
    doff.push_back(ds);
    ds += (size_t)b.nr * b.nc;
  }
  // work items of the T = VH x pass: the low-rank blocks bucketed by rank,
  // each in chunks of kVxCols columns (one workgroup per chunk); a block of
  // several chunks gets partials after the blocks' T (from ts on) and a
  // fixed-order reduction into its T
  std::vector<int2> vx, red;
  std::vector<size_t> vout, rbase;
  size_t ps = ts;
  for (int g = 0; g < 4; g++) {
    c->hm_vx_start[g] = (int)vx.size();
    const int lo = g == 0 ? 1 : (4 << g) + 1, hi = 8 << g;
    for (int a = 0; a < na; a++) {
      if (rank[a] < lo || rank[a] > hi) continue;
      const int nch = (adesc[a].w + hamsx::kVxCols - 1) / hamsx::kVxCols;
      const size_t len = (size_t)rank[a] * nrhs;
      if (nch > 1) {
        red.push_back(make_int2(a, nch));
        rbase.push_back(ps);
      }
      for (int ch = 0; ch < nch; ch++) {
        vx.push_back(make_int2(a, ch));
        vout.push_back(nch > 1 ? ps + (size_t)ch * len : toff[a]);
      }
      if (nch > 1) ps += (size_t)nch * len;
    }
  }
  c->hm_vx_start[4] = (int)vx.size();
  c->hm_vx_nred = (int)red.size();
  if ((rc = c->hmVxItem.ensure(std::max<size_t>(vx.size(), 1))) ||
      (rc = c->hmVxOut.ensure(std::max<size_t>(vx.size(), 1))) ||
      (rc = c->hmVxRed.ensure(std::max<size_t>(red.size(), 1))) ||
      (rc = c->hmVxRbase.ensure(std::max<size_t>(red.size(), 1))))
    return rc;
  if (!vx.empty()) {
    HX_TRY(hipMemcpyAsync(c->hmVxItem.p, vx.data(), sizeof(int2) * vx.size(), hipMemcpyHostToDevice, c->stream));
    HX_TRY(hipMemcpyAsync(c->hmVxOut.p, vout.data(), sizeof(size_t) * vout.size(), hipMemcpyHostToDevice,
                          c->stream));
  }
  if (!red.empty()) {
    HX_TRY(hipMemcpyAsync(c->hmVxRed.p, red.data(), sizeof(int2) * red.size(), hipMemcpyHostToDevice, c->stream));
    HX_TRY(hipMemcpyAsync(c->hmVxRbase.p, rbase.data(), sizeof(size_t) * rbase.size(), hipMemcpyHostToDevice,
                          c->stream));
  }
  const int nd = (int)ddesc.size();
  if ((rc = c->hmDdesc.ensure(std::max(nd, 1))) || (rc = c->hmDoff.ensure(std::max(nd, 1))) ||
      (rc = c->hmD.ensure(std::max<size_t>(ds, 1))) || (rc = c->hmT.ensure(std::max<size_t>(ps, 1))) ||
      (rc = c->hmXp.ensure((size_t)n * nrhs)) || (rc = c->hmYp.ensure((size_t)n * nrhs)))
    return rc;
  if (na > 0) {
    HX_TRY(hipMemcpyAsync(c->hmRank.p, rank.data(), sizeof(int) * na, hipMemcpyHostToDevice, c->stream));
    HX_TRY(hipMemcpyAsync(c->hmToff.p, toff.data(), sizeof(size_t) * na, hipMemcpyHostToDevice, c->stream));
  }
  if (nd > 0) {
    HX_TRY(hipMemcpyAsync(c->hmDdesc.p, ddesc.data(), sizeof(int4) * nd, hipMemcpyHostToDevice, c->stream));
    HX_TRY(hipMemcpyAsync(c->hmDoff.p, doff.data(), sizeof(size_t) * nd, hipMemcpyHostToDevice, c->stream));
  }
  hamsx::launch::hm_dense_gather(c->stream, Aop, n, c->hmPerm.p, nd, c->hmDdesc.p, c->hmDoff.p, c->hmD.p);
  // per-leaf lists, blocks in partition order
  std::vector<int> lstart;
  for (const auto& t : c->hm_tree)
    if (t.left < 0) lstart.push_back(t.s);
  std::sort(lstart.begin(), lstart.end());
  std::vector<std::vector<int>> lists(c->hm_nleaf);
  for (size_t k = 0; k < c->hm_blocks.size(); k++) {
    const HBlk& b = c->hm_blocks[k];
    int q = (int)(std::lower_bound(lstart.begin(), lstart.end(), b.rs) - lstart.begin());
    for (; q < c->hm_nleaf && lstart[q] < b.rs + b.nr; q++) lists[q].push_back(kind[k]);
  }
  std::vector<int> lptr(c->hm_nleaf + 1, 0), litem;
  for (int q = 0; q < c->hm_nleaf; q++) {
    litem.insert(litem.end(), lists[q].begin(), lists[q].end());
    lptr[q + 1] = (int)litem.size();
  }
  // segments of at most kLeafSeg blocks per leaf (one workgroup each); a
  // leaf of several segments gets partial rows and a fixed-order reduction
  std::vector<int4> segs;
  std::vector<size_t> sout, lrbase;
  std::vector<int2> lred;
  size_t yps = 0;
  for (int q = 0; q < c->hm_nleaf; q++) {
    const int ib = lptr[q], ie = lptr[q + 1];
    const int nsg = std::max(1, (ie - ib + hamsx::kLeafSeg - 1) / hamsx::kLeafSeg);
    const size_t len = (size_t)(c->hm_leaves_h[q].y - c->hm_leaves_h[q].x) * nrhs;
    if (nsg > 1) {
      lred.push_back(make_int2(q, nsg));
      lrbase.push_back(yps);
    }
    for (int sgi = 0; sgi < nsg; sgi++) {
      segs.push_back(make_int4(q, ib + sgi * hamsx::kLeafSeg, std::min(ie, ib + (sgi + 1) * hamsx::kLeafSeg),
                               nsg > 1 ? 1 : 0));
      sout.push_back(nsg > 1 ? yps + (size_t)sgi * len : 0);
    }
    if (nsg > 1) yps += (size_t)nsg * len;
  }
  c->hm_nseg = (int)segs.size();
  c->hm_leaf_nred = (int)lred.size();
  if ((rc = c->hmSeg.ensure(std::max<size_t>(segs.size(), 1))) ||
      (rc = c->hmSegOut.ensure(std::max<size_t>(sout.size(), 1))) ||
      (rc = c->hmLeafRed.ensure(std::max<size_t>(lred.size(), 1))) ||
      (rc = c->hmLeafRbase.ensure(std::max<size_t>(lrbase.size(), 1))) ||
      (rc = c->hmYpart.ensure(std::max<size_t>(yps, 1))))
    return rc;
  if (!segs.empty()) {
    HX_TRY(hipMemcpyAsync(c->hmSeg.p, segs.data(), sizeof(int4) * segs.size(), hipMemcpyHostToDevice, c->stream));
    HX_TRY(hipMemcpyAsync(c->hmSegOut.p, sout.data(), sizeof(size_t) * sout.size(), hipMemcpyHostToDevice,
                          c->stream));
  }
  if (!lred.empty()) {
    HX_TRY(hipMemcpyAsync(c->hmLeafRed.p, lred.data(), sizeof(int2) * lred.size(), hipMemcpyHostToDevice,
                          c->stream));
    HX_TRY(hipMemcpyAsync(c->hmLeafRbase.p, lrbase.data(), sizeof(size_t) * lrbase.size(), hipMemcpyHostToDevice,
                          c->stream));
  }
  if ((rc = c->hmLptr.ensure(lptr.size())) || (rc = c->hmLitem.ensure(std::max<size_t>(litem.size(), 1)))) return rc;
  HX_TRY(hipMemcpyAsync(c->hmLptr.p, lptr.data(), sizeof(int) * lptr.size(), hipMemcpyHostToDevice, c->stream));
  if (!litem.empty())
    HX_TRY(hipMemcpyAsync(c->hmLitem.p, litem.data(), sizeof(int) * litem.size(), hipMemcpyHostToDevice,
                          c->stream));
  // mixed precision: complex64 copies of the blocks for the Arnoldi products
  c->hm_c64 = false;
  if (c->prec == HAMSX_PREC_MIXED) {
    if ((rc = c->hmU32.ensure(c->hmU.n)) || (rc = c->hmVH32.ensure(c->hmVH.n)) || (rc = c->hmD32.ensure(c->hmD.n)))
      return rc;
    hamsx::launch::to_c64(c->stream, c->hmU.p, c->hmU32.p, c->hmU.n);
    hamsx::launch::to_c64(c->stream, c->hmVH.p, c->hmVH32.p, c->hmVH.n);
    hamsx::launch::to_c64(c->stream, c->hmD.p, c->hmD32.p, c->hmD.n);
    c->hm_c64 = true;
  }
  HX_TRY(hipStreamSynchronize(c->stream));  // the host vectors above go out of scope
  c->hm_nlr = nlr;
  c->hm_ndense = nd;
  c->hm_nrhs = nrhs;
  c->hm_mean_rank = nlr > 0 ? (double)rank_sum / nlr : 0.0;
  c->hm_ratio = (double)n * n / (double)std::max<size_t>(ds + lr_words, 1);
  c->hm_built = true;
  return check_launch();
}

hamsx::HmDev hm_dev(hamsx_ctx* c) {
  hamsx::HmDev h{};
  h.perm = c->hmPerm.p;
  h.nlr = (int)c->hm_aca_idx.size();
  h.nleaf = c->hm_nleaf;
  h.vx_item = c->hmVxItem.p;
  h.vx_out = c->hmVxOut.p;
  for (int g = 0; g < 5; g++) h.vx_start[g] = c->hm_vx_start[g];
  h.vx_red = c->hmVxRed.p;
  h.vx_rbase = c->hmVxRbase.p;
  h.vx_nred = c->hm_vx_nred;
  h.ldesc = c->hmAdesc.p;
  h.rank = c->hmRank.p;
  h.uoff = c->hmUoff.p;
  h.voff = c->hmVoff.p;
  h.toff = c->hmToff.p;
  h.U = c->hmU.p;
  h.VH = c->hmVH.p;
  h.T = c->hmT.p;
  h.ddesc = c->hmDdesc.p;
  h.doff = c->hmDoff.p;
  h.D = c->hmD.p;
  h.leaves = c->hmLeaves.p;
  h.lptr = c->hmLptr.p;
  h.litem = c->hmLitem.p;
  h.seg = c->hmSeg.p;
  h.seg_out = c->hmSegOut.p;
  h.nseg = c->hm_nseg;
  h.Ypart = c->hmYpart.p;
  h.leaf_red = c->hmLeafRed.p;
  h.leaf_rbase = c->hmLeafRbase.p;
  h.leaf_nred = c->hm_leaf_nred;
  if (c->hm_c64) {
    h.U32 = c->hmU32.p;
    h.VH32 = c->hmVH32.p;
    h.D32 = c->hmD32.p;
  }
  return h;
}

// W = B - A^T (A X) in FP64 from the assembled A (N + L rows): the residual
// of the normal equations when C itself was formed in complex64.
int normal_residual(hamsx_ctx* c, const double2* X, int nrhs, const double2* B, double2* W) {
  int rc;
  if ((rc = c->nY.ensure((size_t)c->nf * nrhs)) || (rc = blas_handle(c))) return rc;
  const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0), mone(-1.0, 0.0);
  HX_TRY(hipMemcpyAsync(W, B, sizeof(double2) * (size_t)c->n * nrhs, hipMemcpyDeviceToDevice, c->stream));
  if (rocblas_zgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c->nf, nrhs, c->n, &one,
                    reinterpret_cast<const rocblas_double_complex*>(c->A.p), c->nf,
                    reinterpret_cast<const rocblas_double_complex*>(X), c->n, &zero,
                    reinterpret_cast<rocblas_double_complex*>(c->nY.p), c->nf) != rocblas_status_success ||
      rocblas_zgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, c->n, nrhs, c->nf, &mone,
                    reinterpret_cast<const rocblas_double_complex*>(c->A.p), c->nf,
                    reinterpret_cast<const rocblas_double_complex*>(c->nY.p), c->nf, &one,
                    reinterpret_cast<rocblas_double_complex*>(W), c->n) != rocblas_status_success)
    return fail(HAMSX_EDEVICE, "rocblas_zgemm (B - A^T A X) failed");
  return 0;
}

// the GMRES operator: Y = beta B + alpha Op X, Op the H-matrix when built,
// else the dense matrix.  inner: a product inside an Arnoldi cycle, which in
// mixed precision reads the complex64 blocks; the residuals at each restart
// and the final one use the FP64 blocks (or A itself, normal_residual).
void op_matvec(hamsx_ctx* c, const double2* Aop, const double2* X, int nrhs, const double2* B, double alpha,
               double beta, double2* Y, bool inner = false) {
  if (c->hm_built)
    hamsx::launch::hm_matvec(c->stream, hm_dev(c), c->n, X, nrhs, B, alpha, beta, Y, c->hmXp.p, c->hmYp.p,
                             inner && c->hm_c64);
  else
    hamsx::launch::matvec(c->stream, Aop, c->n, X, nrhs, B, alpha, beta, Y, c->mvw.p);
}

// X <- Aop^{-1} X; the iteration count adds to last_iters, the final relative
// residual raises last_resid
int do_solve_gmres(hamsx_ctx* c, const double2* Aop, double2* X) {
  const int n = c->n, nr = c->nrhs(), m = c->gm_restart;
  const size_t blk = (size_t)n * nr;
  int rc;
  if ((rc = c->gV.ensure(blk * (m + 1))) || (rc = c->gW.ensure(blk)) || (rc = c->gZ.ensure(blk)) ||
      (rc = c->gU.ensure(blk)) || (rc = c->gB.ensure(blk)) || (rc = c->gDots.ensure((size_t)2 * (m + 1) * nr)) ||
      (rc = c->gNorm.ensure(nr)) || (rc = c->gScale.ensure(nr)) ||
      (rc = c->mvw.ensure(hamsx::launch::matvec_workspace(n, nr))))
    return rc;
  if ((rc = c->gP.ensure(blk)) || (rc = c->gPt.ensure(blk))) return rc;
  if ((rc = c->gH.ensure((size_t)(m + 1) * m * nr)) || (rc = c->gG.ensure((size_t)(m + 1) * nr)) ||
      (rc = c->gSn.ensure((size_t)m * nr)) || (rc = c->gCs.ensure((size_t)m * nr)) || (rc = c->gBn.ensure(nr)) ||
      (rc = c->gWorst.ensure(1)) || (rc = c->gAct.ensure(nr)) || (rc = c->gDone.ensure(nr)) ||
      (rc = c->gMact.ensure(nr)) || (rc = c->gIters.ensure(1)) || (rc = c->gCnt.ensure((size_t)m + 1)))
    return rc;
  // pinned copies of the device counters: [0] unconverged columns at the
  // last restart test, [1 + j] active columns after Arnoldi step j
  const size_t pin_len = (size_t)m + 1;
  if (c->gm_pin_len < pin_len) {
    HX_TRY(hipStreamSynchronize(c->stream));
    if (c->gm_pin) (void)hipHostFree(c->gm_pin);
    c->gm_pin = nullptr;
    c->gm_pin_len = 0;
    HX_TRY(hipHostMalloc((void**)&c->gm_pin, sizeof(int) * pin_len, hipHostMallocDefault));
    c->gm_pin_len = pin_len;
  }
  for (hipEvent_t& e : c->gm_ev)
    if (!e) HX_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if ((rc = gmres_precond_setup(c, Aop))) return rc;
  c->hm_built = false;
  if (c->hm_on && (rc = hm_build(c, Aop, nr))) return rc;
  const hamsx::GmDev gd{c->gH.p, c->gG.p, c->gSn.p, c->gCs.p, c->gBn.p, c->gAct.p, c->gDone.p, c->gMact.p,
                        c->gIters.p};
  // C formed in complex64: the residuals come from A itself (system 1 only)
  const bool exact_normal = c->c_fp32 && Aop == c->Cm.p;
  volatile int* pin = c->gm_pin;
  // B = right-hand sides, X = 0
  HX_TRY(hipMemcpyAsync(c->gB.p, X, sizeof(double2) * blk, hipMemcpyDeviceToDevice, c->stream));
  HX_TRY(hipMemsetAsync(X, 0, sizeof(double2) * blk, c->stream));
  hamsx::launch::col_norm2(c->stream, c->gB.p, n, nr, c->gNorm.p);
  hamsx::launch::gm_start(c->stream, gd, nr, c->gNorm.p);
  double2* d1 = c->gDots.p;
  double2* d2 = c->gDots.p + (size_t)(m + 1) * nr;
  int iter = 0;
  while (iter < c->gm_maxiter) {
    // R = B - A X (into gW) and beta = ||R||; the restart test runs on the
    // device and the host waits once per cycle for the count of columns left
    if (iter == 0)
      HX_TRY(hipMemcpyAsync(c->gW.p, c->gB.p, sizeof(double2) * blk, hipMemcpyDeviceToDevice, c->stream));
    else if (exact_normal) {
      if ((rc = normal_residual(c, X, nr, c->gB.p, c->gW.p))) return rc;
    } else
      op_matvec(c, Aop, X, nr, c->gB.p, -1.0, 1.0, c->gW.p);
    hamsx::launch::col_norm2(c->stream, c->gW.p, n, nr, c->gNorm.p);
    hamsx::launch::gm_restart(c->stream, gd, nr, m, c->gNorm.p, c->gm_tol, c->gScale.p, c->gCnt.p, c->gWorst.p);
    if ((rc = check_launch())) return rc;
    HX_TRY(hipMemcpyAsync((void*)pin, c->gCnt.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HX_TRY(hipStreamSynchronize(c->stream));
    if (pin[0] == 0) break;
    hamsx::launch::col_scale(c->stream, c->gW.p, n, nr, c->gScale.p, c->gV.p);
    // Arnoldi steps, enqueued kGmLag steps ahead of the host's look at the
    // device's count of active columns; a step after every column has
    // converged changes nothing (its scale is 0 and k_gm_step skips the
    // columns), so the cycle ends at most kGmLag steps late.
    int j = 0;
    bool over = false;
    for (; j < m && iter < c->gm_maxiter && !over; j++) {
      iter++;
      double2* Vj = c->gV.p + blk * j;
      gmres_precond_apply(c, Vj, c->gZ.p, nr);
      op_matvec(c, Aop, c->gZ.p, nr, nullptr, 1.0, 0.0, c->gW.p, true);
      hamsx::launch::block_dots(c->stream, c->gV.p, j + 1, c->gW.p, n, nr, d1);
      hamsx::launch::block_update(c->stream, c->gV.p, j + 1, c->gW.p, n, nr, d1, nullptr);
      hamsx::launch::block_dots(c->stream, c->gV.p, j + 1, c->gW.p, n, nr, d2);
      hamsx::launch::block_update(c->stream, c->gV.p, j + 1, c->gW.p, n, nr, d2, nullptr);
      hamsx::launch::col_norm2(c->stream, c->gW.p, n, nr, c->gNorm.p);
      hamsx::launch::gm_step(c->stream, gd, nr, m, j, d1, d2, c->gNorm.p, c->gm_tol, c->gScale.p, c->gCnt.p);
      hamsx::launch::col_scale(c->stream, c->gW.p, n, nr, c->gScale.p, c->gV.p + blk * (j + 1));
      if ((rc = check_launch())) return rc;
      HX_TRY(hipMemcpyAsync((void*)(pin + 1 + j), c->gCnt.p + 1 + j, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HX_TRY(hipEventRecord(c->gm_ev[j % (kGmLag + 1)], c->stream));
      if (j >= kGmLag) {
        HX_TRY(hipEventSynchronize(c->gm_ev[(j - kGmLag) % (kGmLag + 1)]));
        over = pin[1 + j - kGmLag] == 0;
        if (over) iter -= kGmLag;  // the steps after convergence do not count
      }
    }
    // y = H^-1 g per right-hand side (device); X += M^-1 (V y)
    hamsx::launch::gm_solve(c->stream, gd, nr, m, c->gDots.p);
    hamsx::launch::block_combine(c->stream, c->gV.p, j, n, nr, c->gDots.p, c->gU.p);
    gmres_precond_apply(c, c->gU.p, c->gZ.p, nr);
    hamsx::launch::axpy_update(c->stream, X, c->gZ.p, blk);
  }
  // final residual of the operator GMRES worked with, for the statistics
  if (exact_normal) {
    if ((rc = normal_residual(c, X, nr, c->gB.p, c->gW.p))) return rc;
  } else
    op_matvec(c, Aop, X, nr, c->gB.p, -1.0, 1.0, c->gW.p);
  hamsx::launch::col_norm2(c->stream, c->gW.p, n, nr, c->gNorm.p);
  std::vector<double> beta(nr), bnorm(nr);
  int iters = 0;
  HX_TRY(hipMemcpyAsync(beta.data(), c->gNorm.p, sizeof(double) * nr, hipMemcpyDeviceToHost, c->stream));
  HX_TRY(hipMemcpyAsync(bnorm.data(), c->gBn.p, sizeof(double) * nr, hipMemcpyDeviceToHost, c->stream));
  HX_TRY(hipMemcpyAsync(&iters, c->gIters.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HX_TRY(hipStreamSynchronize(c->stream));
  double worst = 0.0;
  for (int q = 0; q < nr; q++)
    if (bnorm[q] >= 1.0e-30) worst = std::max(worst, beta[q] / bnorm[q]);
  c->last_iters += iters;
  c->last_resid = std::max(c->last_resid, worst);
  return check_launch();
}
}  // namespace

// Mixed precision: X <- M^{-1} X by the complex64 factors Af, then FP64
// residual / complex64 correction steps until the residual test passes; the
// test runs on the device and the steps after convergence are launched as
// no-ops, so nothing waits on the host.  state: [converged, steps]; norms:
// the residual and solution norms of the last test.
int refine_solve(hamsx_ctx* c, const double2* M, const float2* Af, const int* perm, double2* X, int* state,
                 double* norms) {
  const int n = c->n, nr = c->nrhs();
  const size_t cnt = (size_t)n * nr;
  hamsx::launch::refine_reset(c->stream, state);
  HX_TRY(hipMemcpyAsync(c->B.p, X, sizeof(double2) * cnt, hipMemcpyDeviceToDevice, c->stream));
  hamsx::launch::lu_solve_c64(c->stream, Af, n, perm, X, c->Xtmp.p, c->trinv.p, nr);
  for (int it = 0; it < c->refine_iters; it++) {
    hamsx::launch::residual(c->stream, M, n, X, c->B.p, c->D.p, nr, norms, c->mvw.p, state);
    hamsx::launch::refine_check(c->stream, norms, nr, c->refine_tol, state);
    hamsx::launch::lu_solve_c64(c->stream, Af, n, perm, c->D.p, c->Xtmp.p, c->trinv.p, nr, state);
    hamsx::launch::axpy_update(c->stream, X, c->D.p, cnt, state);
  }
  // residual of the final iterate (a no-op when the loop converged: the
  // norms then hold the residual that passed the test)
  hamsx::launch::residual(c->stream, M, n, X, c->B.p, c->R.p, nr, norms, c->mvw.p, state);
  return 0;
}

// X <- A^{-1} X for all 6nb + nbeta columns (with refinement in mixed mode)
int solve_factored(hamsx_ctx* c);

int do_solve(hamsx_ctx* c) {
  if (c->have_X) return 0;
  int rc = do_rhs(c);
  if (rc) return rc;
  if (c->solver == HAMSX_SOLVER_GMRES) {
    StageTimer tm(c, 3);
    if ((rc = ensure_normal_matrix(c))) return rc;
    c->last_iters = 0;
    c->last_resid = 0.0;
    if ((rc = do_solve_gmres(c, system_matrix(c), c->X.p))) return rc;
    if (c->sym()) {  // the second system, then MXPOT of both halves
      if ((rc = do_solve_gmres(c, system_matrix2(c), c->X2.p))) return rc;
      hamsx::launch::sym_recombine(c->stream, c->X.p, c->X2.p, (size_t)c->n * c->nrhs());
      if ((rc = check_launch())) return rc;
    }
    c->have_X = true;
    return 0;
  }
  if ((rc = do_factor(c))) return rc;
  return solve_factored(c);
}

// The triangular solves (FP64) or the refined complex64 solves (mixed) of
// the factored system, into X.
int solve_factored(hamsx_ctx* c) {
  int rc = 0;
  const int n = c->n, nr = c->nrhs();
  StageTimer tm(c, 3);
  if (c->prec == HAMSX_PREC_FP64) {
    const bool two = c->sym() && sym_second_stream(c) != nullptr;
    if (two) {  // system 2's solves on its own stream, at once with system 1's
      if ((rc = c->Xtmp2.ensure((size_t)n * nr))) return rc;
      HX_TRY(hipEventRecord(c->sym_go, c->stream));
      HX_TRY(hipStreamWaitEvent(c->sym_stream, c->sym_go, 0));
      if ((rc = c->trinv2.ensure(hamsx::launch::tri_inverse_len(n)))) return rc;
      hamsx::launch::lu_solve(c->sym_stream, system_matrix2(c), n, c->perm2.p, c->X2.p, c->Xtmp2.p, c->trinv2.p, nr);
      HX_TRY(hipEventRecord(c->sym_done, c->sym_stream));
    }
    hamsx::launch::lu_solve(c->stream, system_matrix(c), n, c->perm.p, c->X.p, c->Xtmp.p, c->trinv.p, nr);
    if (two) HX_TRY(hipStreamWaitEvent(c->stream, c->sym_done, 0));
    else if (c->sym())
      hamsx::launch::lu_solve(c->stream, system_matrix2(c), n, c->perm2.p, c->X2.p, c->Xtmp.p, c->trinv.p, nr);
  } else {
    if ((rc = refine_solve(c, system_matrix(c), c->Af.p, c->perm.p, c->X.p, c->flags.p + 2, c->norms.p))) return rc;
    if (c->sym() &&
        (rc = refine_solve(c, system_matrix2(c), c->Af2.p, c->perm2.p, c->X2.p, c->flags.p + 6, c->norms2.p)))
      return rc;
  }
  if (c->sym()) hamsx::launch::sym_recombine(c->stream, c->X.p, c->X2.p, (size_t)n * nr);
  c->have_X = true;
  return check_launch();
}

int fetch_mxpot(hamsx_ctx* c, int col0, int ncol, double* dst) {
  if (!dst || ncol <= 0) return 0;
  const size_t len = (size_t)c->n * ncol;
  HX_TRY(hipMemcpyAsync(dst, c->X.p + (size_t)c->n * col0, sizeof(double2) * len, hipMemcpyDeviceToHost, c->stream));
  if (c->sym())  // MXPOT(:, :, 2): the mirror half
    HX_TRY(hipMemcpyAsync(dst + 2 * len, c->X2.p + (size_t)c->n * col0, sizeof(double2) * len,
                          hipMemcpyDeviceToHost, c->stream));
  HX_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

// Enqueues the force reductions and the copies of results, flags and
// refinement norms into the pinned staging area (no host wait).
int enqueue_forces(hamsx_ctx* c) {
  const size_t len = c->out_len();
  // staging layout: [wvn: NK][v, wk, w1, h][beta: nbeta][results: len][flags: 8 ints][norms: 2 * nrhs]
  // [norms2: 2 * nrhs]
  int rc = ensure_pinned(c, kStageHead + c->nbeta + len + 4 + 4 * (size_t)c->nrhs());
  if (rc) return rc;
  double* hout = c->h_pinned + kStageHead + c->nbeta;
  {
    StageTimer tm(c, 4);
    hamsx::launch::forces(c->stream, c->mesh(), c->freq, c->waves(), c->nbody, c->body_start_d.p, c->X.p,
                          c->out.p, c->sym() ? c->X2.p : nullptr, c->sym_mirror);
  }
  if ((rc = check_launch())) return rc;
  HX_TRY(hipMemcpyAsync(hout, c->out.p, sizeof(double) * len, hipMemcpyDeviceToHost, c->stream));
  HX_TRY(hipMemcpyAsync(hout + len, c->flags.p, sizeof(int) * 8, hipMemcpyDeviceToHost, c->stream));
  if (c->prec == HAMSX_PREC_MIXED && c->nrhs() > 0 && c->sym() && c->solver == HAMSX_SOLVER_LU)
    HX_TRY(hipMemcpyAsync(hout + len + 4 + 2 * (size_t)c->nrhs(), c->norms2.p, sizeof(double) * 2 * c->nrhs(),
                          hipMemcpyDeviceToHost, c->stream));
  if (c->prec == HAMSX_PREC_MIXED && c->nrhs() > 0)
    HX_TRY(hipMemcpyAsync(hout + len + 4, c->norms.p, sizeof(double) * 2 * c->nrhs(), hipMemcpyDeviceToHost,
                          c->stream));
  return 0;
}

// Waits for the stream and unpacks the staging area written by enqueue_forces.
int complete_forces(hamsx_ctx* c, double* amas, double* bdmp, double* exfc, int* info_out) {
  const int nd = 6 * c->nbody;
  const size_t len = c->out_len();
  const double* hout = c->h_pinned + kStageHead + c->nbeta;
  HX_TRY(hipStreamSynchronize(c->stream));
  harvest_marks(c);
  if (amas) std::memcpy(amas, hout, sizeof(double) * nd * nd);
  if (bdmp) std::memcpy(bdmp, hout + (size_t)nd * nd, sizeof(double) * nd * nd);
  if (exfc) std::memcpy(exfc, hout + 2 * (size_t)nd * nd, sizeof(double) * 2 * nd * c->nbeta);
  // quadrature overflow, INFO, refinement converged, correction steps, -, INFO of symmetry
  // system 2, its refinement converged and steps
  int fl[8];
  std::memcpy(fl, hout + len, sizeof fl);
  if (info_out) *info_out = fl[1] ? fl[1] : (c->sym() ? fl[5] : 0);
  bool noconv = false;
  if (c->prec == HAMSX_PREC_MIXED && c->nrhs() > 0 && c->solver == HAMSX_SOLVER_LU) {
    const int nsys = c->sym() ? 2 : 1, nr = c->nrhs();
    double worst = 0.0;
    c->last_iters = 0;
    for (int ip = 0; ip < nsys; ip++) {
      const double* nm = hout + len + 4 + 2 * (size_t)nr * ip;
      double w = 0.0;
      for (int q = 0; q < nr; q++) {
        double rn = nm[q], xn = nm[nr + q];
        w = std::max(w, xn > 0 ? rn / xn : rn);
      }
      worst = std::max(worst, w);
      c->last_iters += fl[3 + 4 * ip];
      noconv = noconv || (!fl[2 + 4 * ip] && c->refine_tol > 0.0 && w > c->refine_tol);
    }
    c->last_resid = worst;
  }
  if (fl[0] & 1) return fail(HAMSX_EDEVICE, "adaptive quadrature stack exhausted in the finite-depth kernel");
  if (noconv)
    return fail(HAMSX_ENOCONV, "iterative refinement: residual %.3e above %.3e after %d steps", c->last_resid,
                c->refine_tol, c->last_iters);
  return 0;
}

int finish_forces(hamsx_ctx* c, double* amas, double* bdmp, double* exfc, int* info_out) {
  int rc = enqueue_forces(c);
  if (rc) return rc;
  return complete_forces(c, amas, bdmp, exfc, info_out);
}

// Field points per pass: the 16-bit row index of the pair list bounds the
// binned form to n + F <= 65535, and M = [S | T] (F x 2n complex) is kept
// under 4 GiB.  HAMSX_FIELD_CHUNK (diagnostic) lowers the count further, to
// exercise the pass loop on small cases.
int field_chunk(const hamsx_ctx* c, bool binned) {
  size_t cap = ((size_t)4 << 30) / (32 * (size_t)c->n);
  if (binned) cap = std::min(cap, (size_t)(0xFFFF - c->n));
  cap = std::min<size_t>(cap, 32768);
  if (const char* e = std::getenv("HAMSX_FIELD_CHUNK")) {
    const long v = std::atol(e);
    if (v > 0) cap = std::min<size_t>(cap, (size_t)v);
  }
  return (int)std::max<size_t>(1, std::min<size_t>(cap, (size_t)c->nfp));
}

int busy(hamsx_ctx*) {
  return fail(HAMSX_ESTATE, "a submitted frequency solve is pending on this context: call hamsx_collect first");
}

}  // namespace

extern "C" {

const char* hamsx_version(void) { return "hamsx 0.1 (gfx950)"; }

const char* hamsx_last_error(void) { return g_last_error.c_str(); }

int hamsx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int hamsx_create(hamsx_ctx** out, int device) {
  if (!out) return fail(HAMSX_EARG, "ctx pointer is NULL");
  *out = nullptr;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) return fail(HAMSX_EDEVICE, "no HIP device available");
  if (device < 0 || device >= nd) return fail(HAMSX_EARG, "device %d out of range (%d devices)", device, nd);
  auto* c = new hamsx_ctx();
  c->device = device;
  tl_alloc_epoch = &c->alloc_epoch;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(HAMSX_EDEVICE, "stream creation failed on device %d", device);
  }
  // LU look-ahead: a second stream (default priority) and its two events
  if (hipStreamCreateWithFlags(&c->lu_aux.s2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->lu_aux.panel, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->lu_aux.update, hipEventDisableTiming) != hipSuccess) {
    hamsx_destroy(c);
    return fail(HAMSX_EDEVICE, "stream/event creation failed on device %d", device);
  }
  int rc = c->flags.ensure(8);
  if (rc == 0) rc = c->wvn.ensure(kStageHead);  // the roots, then the scalars read by freq_now
  if (rc) {
    hamsx_destroy(c);
    return rc;
  }
  if (hipMemsetAsync(c->flags.p, 0, sizeof(int) * 8, c->stream) != hipSuccess) {
    hamsx_destroy(c);
    return fail(HAMSX_EDEVICE, "flag initialisation failed");
  }
  *out = c;
  return 0;
}

int hamsx_destroy(hamsx_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  if (tl_alloc_epoch == &c->alloc_epoch) tl_alloc_epoch = nullptr;
  for (auto* b : {&c->cx, &c->cy, &c->cz, &c->area, &c->pnsz, &c->nrm, &c->frame, &c->wvn, &c->beta, &c->out,
                  &c->norms})
    b->release();
  for (auto* b : {&c->ncn, &c->body, &c->body_start_d, &c->chunk_body, &c->chunk_j0, &c->chunk_j1, &c->ipiv,
                  &c->perm, &c->widx, &c->wcnt, &c->flags})
    b->release();
  for (auto* b : {&c->fx, &c->fy, &c->fz}) b->release();
  c->Cm.release();
  c->Xe.release();
  if (c->blas) (void)rocblas_destroy_handle(c->blas);
  for (auto* b : {&c->gV, &c->gW, &c->gZ, &c->gU, &c->gB, &c->gP, &c->gPt, &c->gDots, &c->pLU}) b->release();
  c->gNorm.release();
  c->gScale.release();
  c->pPerm.release();
  c->pIpiv.release();
  c->modetab.release();
  for (auto* b : {&c->A, &c->S, &c->X, &c->Xtmp, &c->B, &c->R, &c->D, &c->dum2, &c->partial, &c->ludiag, &c->mvw})
    b->release();
  c->Af.release();
  for (auto& m : c->marks) c->pool.insert(c->pool.end(), {m.a, m.b});
  for (auto e : c->pool) (void)hipEventDestroy(e);
  if (c->h_pinned) (void)hipHostFree(c->h_pinned);
  if (c->gm_pin) (void)hipHostFree(c->gm_pin);
  for (hipEvent_t e : c->gm_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->sym_stream) (void)hipStreamSynchronize(c->sym_stream);
  for (hipEvent_t ev : {c->sym_go, c->sym_done, c->grp_ev})
    if (ev) (void)hipEventDestroy(ev);
  if (c->sym_stream) (void)hipStreamDestroy(c->sym_stream);
  if (c->lu_aux.s2) (void)hipStreamSynchronize(c->lu_aux.s2);
  if (c->lu_aux.panel) (void)hipEventDestroy(c->lu_aux.panel);
  if (c->lu_aux.update) (void)hipEventDestroy(c->lu_aux.update);
  if (c->lu_aux.s2) (void)hipStreamDestroy(c->lu_aux.s2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int hamsx_set_option(hamsx_ctx* c, int option, int value) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  drop_graph(c);
  switch (option) {
    case HAMSX_OPT_LU_LOOKAHEAD:
      if (value < -1 || value > 1) return fail(HAMSX_EARG, "HAMSX_OPT_LU_LOOKAHEAD is -1, 0 or 1");
      c->opt_lookahead = value;
      return 0;
    case HAMSX_OPT_LU_SELECT_ROWS:
      if (value != 256 && value != 512 && value != 1024)
        return fail(HAMSX_EARG, "HAMSX_OPT_LU_SELECT_ROWS is 256, 512 or 1024");
      c->lu_aux.sel_rows = value;
      return 0;
    case HAMSX_OPT_LU_OUTER_MIN_N:
      c->lu_aux.outer_min_n = value;
      c->outer_default = false;
      return 0;
    case HAMSX_OPT_GRAPHS:
      if (value != 0 && value != 1) return fail(HAMSX_EARG, "HAMSX_OPT_GRAPHS is 0 or 1");
      c->opt_graphs = value != 0;
      return 0;
    case HAMSX_OPT_FP32_FARFIELD:
      if (value != 0 && value != 1) return fail(HAMSX_EARG, "HAMSX_OPT_FP32_FARFIELD is 0 or 1");
      c->opt_far32 = value != 0;
      c->have_A = c->have_S = c->have_LU = c->have_X = false;
      return 0;
    default:
      return fail(HAMSX_EARG, "unknown option %d", option);
  }
}

int hamsx_get_option(hamsx_ctx* c, int option, int* value) {
  if (!c || !value) return fail(HAMSX_EARG, "ctx or value is NULL");
  switch (option) {
    case HAMSX_OPT_LU_LOOKAHEAD: *value = c->opt_lookahead; return 0;
    case HAMSX_OPT_LU_SELECT_ROWS: *value = c->lu_aux.sel_rows; return 0;
    case HAMSX_OPT_LU_OUTER_MIN_N: *value = c->lu_aux.outer_min_n; return 0;
    case HAMSX_OPT_GRAPHS: *value = c->opt_graphs ? 1 : 0; return 0;
    case HAMSX_OPT_FP32_FARFIELD: *value = c->opt_far32 ? 1 : 0; return 0;
    default: return fail(HAMSX_EARG, "unknown option %d", option);
  }
}

int hamsx_set_precision(hamsx_ctx* c, int mode) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  drop_graph(c);
  if (mode != HAMSX_PREC_FP64 && mode != HAMSX_PREC_MIXED) return fail(HAMSX_EARG, "unknown precision mode %d", mode);
  if (mode != c->prec) {
    c->prec = mode;
    c->have_LU = false;
    c->have_X = false;
  }
  return 0;
}

int hamsx_panel_geometry(int nnode, const double* xyz, int nelem, const int* ncn, const int* ncon, const double* xr,
                         double* cen, double* area, double* pnsz, double* dxyz) {
  if (nnode <= 0 || nelem <= 0 || !xyz || !ncn || !ncon || !xr || !cen || !area || !pnsz || !dxyz)
    return fail(HAMSX_EARG, "hamsx_panel_geometry: bad arguments");
  for (int e = 0; e < nelem; e++) {
    const int nc = ncn[e];
    if (nc != 3 && nc != 4) return fail(HAMSX_EARG, "panel %d has %d vertices (3 or 4 expected)", e + 1, nc);
    double p[4][3];
    for (int c = 0; c < nc; c++) {
      int nd = ncon[e + (size_t)nelem * c] - 1;
      if (nd < 0 || nd >= nnode) return fail(HAMSX_EARG, "panel %d refers to node %d", e + 1, nd + 1);
      for (int d = 0; d < 3; d++) p[c][d] = xyz[nd + (size_t)nnode * d];
    }
    for (int d = 0; d < 3; d++) {  // CalPanelCentre
      double s = 0.0;
      for (int c = 0; c < nc; c++) s = s + p[c][d];
      cen[e + (size_t)nelem * d] = s / nc;
    }
    auto tri = [](const double* p1, const double* p2, const double* p3) {  // CalDeltaArea_Improved
      double v21[3], v23[3], v13[3];
      for (int d = 0; d < 3; d++) {
        v21[d] = p1[d] - p2[d];
        v23[d] = p3[d] - p2[d];
        v13[d] = p3[d] - p1[d];
      }
      double a = norm3(v21), b = norm3(v13), c = norm3(v23);
      return std::sqrt((a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c))) / 4.0;
    };
    double ds = tri(p[0], p[1], p[2]);
    if (nc == 4) ds = ds + tri(p[0], p[3], p[2]);
    area[e] = ds;
    double mx = 0.0;  // CalPanelChartLength
    for (int c = 0; c < nc; c++) {
      const double* a = p[c];
      const double* b = p[(c + 1) % nc];
      double v[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      double l = norm3(v);
      mx = (c == 0) ? l : std::max(mx, l);
    }
    pnsz[e] = mx;
    double un[3];  // CalTransNormals
    if (nc == 3) {
      double v21[3], v23[3];
      for (int d = 0; d < 3; d++) {
        v21[d] = p[0][d] - p[1][d];
        v23[d] = p[2][d] - p[1][d];
      }
      un[0] = v21[1] * v23[2] - v21[2] * v23[1];
      un[1] = v21[2] * v23[0] - v21[0] * v23[2];
      un[2] = v21[0] * v23[1] - v21[1] * v23[0];
    } else {
      double v13[3], v24[3];
      for (int d = 0; d < 3; d++) {
        v13[d] = p[2][d] - p[0][d];
        v24[d] = p[3][d] - p[1][d];
      }
      un[0] = v24[1] * v13[2] - v24[2] * v13[1];
      un[1] = v24[2] * v13[0] - v24[0] * v13[2];
      un[2] = v24[0] * v13[1] - v24[1] * v13[0];
    }
    double aun = norm3(un);
    if (!(aun > 0.0)) return fail(HAMSX_EARG, "panel %d is degenerate (zero normal)", e + 1);
    double n1 = un[0] / aun, n2 = un[1] / aun, n3 = un[2] / aun;
    double x = cen[e], y = cen[e + (size_t)nelem], z = cen[e + 2 * (size_t)nelem];
    dxyz[e] = n1;
    dxyz[e + (size_t)nelem] = n2;
    dxyz[e + 2 * (size_t)nelem] = n3;
    dxyz[e + 3 * (size_t)nelem] = (y - xr[1]) * n3 - (z - xr[2]) * n2;  // CalRotNormals
    dxyz[e + 4 * (size_t)nelem] = (z - xr[2]) * n1 - (x - xr[0]) * n3;
    dxyz[e + 5 * (size_t)nelem] = (x - xr[0]) * n2 - (y - xr[1]) * n1;
  }
  return 0;
}

int hamsx_wave_properties(int inft, double value, double h, double* v, double* wk, double* w1, double* wvn) {
  if (!v || !wk || !w1) return fail(HAMSX_EARG, "hamsx_wave_properties: NULL output");
  if (wvn) std::fill(wvn, wvn + HAMSX_NK, 0.0);
  if (inft == HAMSX_FREQ_LIMIT) {
    // SYBO = 1 (ImplementSubs.f90:45-70): V = WK = W1 = 0 for the zero-frequency
    // limit, -1 for the infinite-frequency limit; no dispersion roots
    if (value != 0.0 && value != -1.0) return fail(HAMSX_EARG, "a frequency limit is 0 (zero) or -1 (infinite)");
    *v = *wk = *w1 = value;
    return 0;
  }
  if (!(value > 0.0)) return fail(HAMSX_EARG, "frequency value must be positive");
  double tmp[HAMSX_NK] = {0};
  double* roots = wvn ? wvn : tmp;
  const bool deep = h <= 0.0;
  switch (inft) {  // CalWaveProperts, ImplementSubs.f90:72-139
    case 1:
      *v = value;
      *w1 = std::sqrt(kG * *v);
      if (deep) *wk = *v;
      else {
        dispersion(*w1, h, HAMSX_NK, roots);
        *wk = roots[0];
      }
      break;
    case 2:
      *wk = value;
      *v = deep ? *wk : *wk * std::tanh(*wk * h);
      *w1 = std::sqrt(kG * *v);
      if (!deep) dispersion(*w1, h, HAMSX_NK, roots);
      break;
    case 3:
    case 4:
      *w1 = inft == 3 ? value : 2.0 * kPi / value;
      *v = *w1 * *w1 / kG;
      if (deep) *wk = *v;
      else {
        dispersion(*w1, h, HAMSX_NK, roots);
        *wk = roots[0];
      }
      break;
    case 5:
      *wk = 2.0 * kPi / value;
      *v = deep ? *wk : *wk * std::tanh(*wk * h);
      *w1 = std::sqrt(kG * *v);
      if (!deep) dispersion(*w1, h, HAMSX_NK, roots);
      break;
    default:
      return fail(HAMSX_EARG, "input frequency type %d (1..5 expected)", inft);
  }
  return 0;
}

int hamsx_set_mesh(hamsx_ctx* c, int nelem, int nnode, const double* xyz, const int* ncn, const int* ncon,
                   const double* cen, const double* area, const double* pnsz, const double* dxyz, int nbody,
                   const int* body_start) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  drop_graph(c);
  if (nelem <= 0 || nnode <= 0 || nbody <= 0 || !xyz || !ncn || !ncon || !cen || !area || !pnsz || !dxyz ||
      !body_start)
    return fail(HAMSX_EARG, "hamsx_set_mesh: bad arguments");
  if (body_start[0] != 0 || body_start[nbody] != nelem)
    return fail(HAMSX_EARG, "body_start must run from 0 to nelem");
  for (int b = 0; b < nbody; b++)
    if (body_start[b + 1] <= body_start[b]) return fail(HAMSX_EARG, "body %d has no panels", b + 1);
  int rc = set_device(c);
  if (rc) return rc;
  HX_TRY(hipStreamSynchronize(c->stream));
  const size_t n = nelem;
  // Newman frames, body index and right-hand-side chunks on the host
  std::vector<double> frame(n * kFrameLen);
  std::vector<int> body(n);
  for (size_t e = 0; e < n; e++) {
    const int nc = ncn[e];
    if (nc != 3 && nc != 4) return fail(HAMSX_EARG, "panel %zu has %d vertices", e + 1, nc);
    double xv[12];
    for (int q = 0; q < nc; q++) {
      int nd = ncon[e + n * q] - 1;
      if (nd < 0 || nd >= nnode) return fail(HAMSX_EARG, "panel %zu refers to node %d", e + 1, nd + 1);
      for (int d = 0; d < 3; d++) xv[3 * q + d] = xyz[nd + (size_t)nnode * d];
    }
    newman_frame(nc, xv, frame.data() + e * kFrameLen);
  }
  std::vector<double> vtx(n * 12, 0.0);
  for (size_t e = 0; e < n; e++)
    for (int q = 0; q < ncn[e]; q++) {
      const int nd = ncon[e + n * q] - 1;
      for (int d = 0; d < 3; d++) vtx[(size_t)(3 * q + d) * n + e] = xyz[nd + (size_t)nnode * d];
    }
  std::vector<int> cb, c0, c1;
  for (int b = 0; b < nbody; b++) {
    for (int j = body_start[b]; j < body_start[b + 1]; j++) body[j] = b;
    for (int j = body_start[b]; j < body_start[b + 1]; j += 256) {
      cb.push_back(b);
      c0.push_back(j);
      c1.push_back(std::min(j + 256, body_start[b + 1]));
    }
  }
  c->n = nelem;
  c->bins.n = 0;
  c->hm_part = c->hm_built = false;
  c->hcen.assign(cen, cen + 3 * n);
  c->nbody = nbody;
  c->nchunk = (int)cb.size();
  c->body_start.assign(body_start, body_start + nbody + 1);
  if ((rc = c->cx.ensure(n)) || (rc = c->cy.ensure(n)) || (rc = c->cz.ensure(n)) || (rc = c->area.ensure(n)) ||
      (rc = c->pnsz.ensure(n)) || (rc = c->nrm.ensure(6 * n)) || (rc = c->frame.ensure(n * kFrameLen)) ||
      (rc = c->ncn.ensure(n)) || (rc = c->body.ensure(n)) || (rc = c->body_start_d.ensure(nbody + 1)) ||
      (rc = c->chunk_body.ensure(cb.size())) || (rc = c->chunk_j0.ensure(cb.size())) ||
      (rc = c->chunk_j1.ensure(cb.size())) || (rc = c->A.ensure(n * n)) || (rc = c->S.ensure(n * n)))
    return rc;
  HX_TRY(hipMemcpy(c->cx.p, cen, sizeof(double) * n, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->cy.p, cen + n, sizeof(double) * n, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->cz.p, cen + 2 * n, sizeof(double) * n, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->area.p, area, sizeof(double) * n, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->pnsz.p, pnsz, sizeof(double) * n, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->nrm.p, dxyz, sizeof(double) * 6 * n, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->frame.p, frame.data(), sizeof(double) * n * kFrameLen, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->ncn.p, ncn, sizeof(int) * n, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->body.p, body.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->body_start_d.p, body_start, sizeof(int) * (nbody + 1), hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->chunk_body.p, cb.data(), sizeof(int) * cb.size(), hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->chunk_j0.p, c0.data(), sizeof(int) * cb.size(), hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->chunk_j1.p, c1.data(), sizeof(int) * cb.size(), hipMemcpyHostToDevice));
  if ((rc = c->vtx.ensure(n * 12))) return rc;
  HX_TRY(hipMemcpy(c->vtx.p, vtx.data(), sizeof(double) * n * 12, hipMemcpyHostToDevice));
  c->nfp = 0;
  c->bodymat_nbody = 0;
  c->field_gen++;
  c->sym_mirror = -1;
  c->mbins.n = 0;
  c->nlid = 0;
  c->nf = nelem;
  if ((rc = c->fx.ensure(n)) || (rc = c->fy.ensure(n)) || (rc = c->fz.ensure(n))) return rc;
  HX_TRY(hipMemcpy(c->fx.p, cen, sizeof(double) * n, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->fy.p, cen + n, sizeof(double) * n, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->fz.p, cen + 2 * n, sizeof(double) * n, hipMemcpyHostToDevice));
  c->have_A = c->have_S = c->have_LU = c->have_X = c->have_C = false;
  return 0;
}

static int build_mirror_view(hamsx_ctx* c, int mirror);

int hamsx_set_lid(hamsx_ctx* c, int nlid, const double* lid_cen) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  drop_graph(c);
  if (c->n == 0) return fail(HAMSX_ESTATE, "hamsx_set_lid needs hamsx_set_mesh first");
  if (nlid < 0 || (nlid > 0 && !lid_cen)) return fail(HAMSX_EARG, "bad lid arguments");
  if (c->n + nlid > 0xFFFF) return fail(HAMSX_EARG, "hull + lid panels exceed 65535");
  int rc = set_device(c);
  if (rc) return rc;
  HX_TRY(hipStreamSynchronize(c->stream));
  const size_t n = c->n, nf = n + nlid;
  std::vector<double> f(3 * nf);
  for (int d = 0; d < 3; d++) {
    HX_TRY(hipMemcpy(f.data() + d * nf, (d == 0 ? c->cx.p : d == 1 ? c->cy.p : c->cz.p), sizeof(double) * n,
                     hipMemcpyDeviceToHost));
    for (int l = 0; l < nlid; l++) f[d * nf + n + l] = lid_cen[l + (size_t)nlid * d];
  }
  if ((rc = c->fx.ensure(nf)) || (rc = c->fy.ensure(nf)) || (rc = c->fz.ensure(nf))) return rc;
  HX_TRY(hipMemcpy(c->fx.p, f.data(), sizeof(double) * nf, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->fy.p, f.data() + nf, sizeof(double) * nf, hipMemcpyHostToDevice));
  HX_TRY(hipMemcpy(c->fz.p, f.data() + 2 * nf, sizeof(double) * nf, hipMemcpyHostToDevice));
  c->nlid = nlid;
  c->nf = (int)nf;
  c->bins.n = 0;  // new field rows: the pair bins are rebuilt
  if (c->sym() && (rc = build_mirror_view(c, c->sym_mirror))) return rc;
  c->have_A = c->have_S = c->have_LU = c->have_X = c->have_C = false;
  return 0;
}

int hamsx_set_frequency(hamsx_ctx* c, double v, double wk, double w1, double h, const double* wvn, int nk) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  if (h > 0.0 && (!wvn || nk < hamsx::kNK)) return fail(HAMSX_EARG, "finite depth needs %d dispersion roots", hamsx::kNK);
  int rc = set_device(c);
  if (rc) return rc;
  if ((rc = ensure_pinned(c, kStageHead + std::max(1, c->nbeta) + c->out_len() + 4 + 4 * (size_t)c->nrhs())))
    return rc;
  // the staging slot is free: every earlier use ended in a stream synchronisation
  HX_TRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < hamsx::kNK; k++) c->h_pinned[k] = (h > 0.0 && wvn) ? wvn[k] : 0.0;
  c->h_pinned[hamsx::kNK] = v;
  c->h_pinned[hamsx::kNK + 1] = wk;
  c->h_pinned[hamsx::kNK + 2] = w1;
  c->h_pinned[hamsx::kNK + 3] = h;
  HX_TRY(hipMemcpyAsync(c->wvn.p, c->h_pinned, sizeof(double) * kStageHead, hipMemcpyHostToDevice, c->stream));
  c->freq = hamsx::FreqDev{v, wk, w1, h, c->wvn.p};
  c->freq.sc = c->wvn.p + hamsx::kNK;
  c->have_freq = true;
  c->have_A = c->have_S = c->have_LU = c->have_X = false;
  return 0;
}

int hamsx_set_waves(hamsx_ctx* c, int nbeta, const double* beta, double amp, double xw0, double xw1, double rho,
                    int isol) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  drop_graph(c);
  if (nbeta < 0 || (nbeta > 0 && !beta)) return fail(HAMSX_EARG, "bad heading list");
  if (isol != 1 && isol != 2) return fail(HAMSX_EARG, "ISOL must be 1 or 2");
  int rc = set_device(c);
  if (rc) return rc;
  HX_TRY(hipStreamSynchronize(c->stream));
  if ((rc = c->beta.ensure(std::max(1, nbeta)))) return rc;
  if (nbeta > 0) HX_TRY(hipMemcpy(c->beta.p, beta, sizeof(double) * nbeta, hipMemcpyHostToDevice));
  c->nbeta = nbeta;
  c->amp = amp;
  c->xw0 = xw0;
  c->xw1 = xw1;
  c->rho = rho;
  c->isol = isol;
  c->have_waves = true;
  c->have_X = false;
  return 0;
}

int hamsx_assemble(hamsx_ctx* c) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  if (c->n == 0) return fail(HAMSX_ESTATE, "hamsx_assemble needs hamsx_set_mesh first");
  if (!c->have_freq) return fail(HAMSX_ESTATE, "hamsx_assemble needs hamsx_set_frequency first");
  int rc = set_device(c);
  if (rc) return rc;
  const bool finite = c->freq.h >= 0.0;
  const bool binned = finite && c->nf <= 0xFFFF && !force_grid_assembly();
  if ((rc = c->A.ensure((size_t)c->nf * c->n)) || (rc = c->S.ensure((size_t)c->nf * c->n))) return rc;
  if (binned && (rc = ensure_bins(c, c->bins, c->mesh(), 0))) return rc;
  if (finite && (rc = c->modetab.ensure(hamsx::mode_tab_len(c->nf)))) return rc;
  c->have_C = false;
  c->freq.tab = finite ? c->modetab.p : nullptr;
  c->freq.far32 = c->prec == HAMSX_PREC_MIXED && c->opt_far32 ? 1 : 0;
  c->freq.limes_ieee = force_limes_ieee() ? 1 : 0;
  const bool limit = c->freq.v == 0.0 || c->freq.v == -1.0;
  {
    StageTimer tm(c, 0);
    if (finite && !limit) hamsx::launch::mode_tables(c->stream, c->mesh(), c->freq);
    if (binned)
      assemble_binned(c, c->bins, c->mesh(), c->freq, c->A.p, c->S.p, c->flags.p, true);
    else
      hamsx::launch::assemble(c->stream, c->mesh(), c->freq, c->A.p, c->S.p, c->flags.p);
  }
  if ((rc = check_launch())) return rc;
  if (c->sym()) {
    // the mirrored field points (IS = 2), then A +- T2 and S +- S2
    const size_t nn = (size_t)c->nf * c->n;
    const hamsx::MeshDev mv = c->mirror_view();
    if ((rc = c->T2.ensure(nn)) || (rc = c->S2.ensure(nn)) ||
        (finite && (rc = c->mmodetab.ensure(hamsx::mode_tab_len(mv.nf)))))
      return rc;
    const bool mbinned = finite && mv.nf <= 0xFFFF && !force_grid_assembly();
    if (mbinned && (rc = ensure_bins(c, c->mbins, mv, 1 + (unsigned long long)c->sym_mirror + 2 * (unsigned long long)c->nlid)))
      return rc;
    hamsx::FreqDev fm = c->freq;
    fm.tab = finite ? c->mmodetab.p : nullptr;
    StageTimer tm(c, 0);
    if (finite && !limit) hamsx::launch::mode_tables(c->stream, mv, fm);
    if (mbinned)
      assemble_binned(c, c->mbins, mv, fm, c->T2.p, c->S2.p, c->flags.p, false);
    else
      hamsx::launch::assemble(c->stream, mv, fm, c->T2.p, c->S2.p, c->flags.p);
    hamsx::launch::sym_combine(c->stream, c->A.p, c->T2.p, c->S.p, c->S2.p, nn);
    if ((rc = check_launch())) return rc;
  }
  c->have_A = c->have_S = true;
  c->have_LU = c->have_X = false;
  return 0;
}

// The mirror view's points for plane `mirror` from the main view's field
// points (hull centroids, then lid points): rows n .. n + nf - 1 evaluate at
// the mirrored points; FLAG is taken at the unmirrored hull centroid
// (CalGreenFuncMulti.f90:134-139) and at the mirrored lid point (the
// distance CALGREEN_MULTI_IRR gives the Green function, :278-284).
static int build_mirror_view(hamsx_ctx* c, int mirror) {
  const size_t n = c->n, nf = c->nf, m2 = n + nf;
  std::vector<double> f(3 * nf), h(6 * m2);
  for (int d = 0; d < 3; d++)
    HX_TRY(hipMemcpy(f.data() + d * nf, (d == 0 ? c->fx.p : d == 1 ? c->fy.p : c->fz.p), sizeof(double) * nf,
                     hipMemcpyDeviceToHost));
  for (int d = 0; d < 3; d++) {
    const double sg = d == mirror ? -1.0 : 1.0;
    for (size_t i = 0; i < n; i++) h[d * m2 + i] = h[(3 + d) * m2 + i] = f[d * nf + i];
    for (size_t i = 0; i < nf; i++) {
      const double v = f[d * nf + i];
      h[d * m2 + n + i] = sg * v;
      h[(3 + d) * m2 + n + i] = i < n ? v : sg * v;
    }
  }
  int rc;
  if ((rc = c->mview.ensure(6 * m2))) return rc;
  HX_TRY(hipMemcpy(c->mview.p, h.data(), sizeof(double) * 6 * m2, hipMemcpyHostToDevice));
  c->mbins.n = 0;
  return 0;
}

int hamsx_set_symmetry(hamsx_ctx* c, int isx, int isy) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  drop_graph(c);
  if (c->n == 0) return fail(HAMSX_ESTATE, "hamsx_set_symmetry needs hamsx_set_mesh first");
  if ((isx != 0 && isx != 1) || (isy != 0 && isy != 1)) return fail(HAMSX_EARG, "ISX and ISY must be 0 or 1");
  if (isx && isy) return fail(HAMSX_EARG, "ISX and ISY cannot be simultaneously 1");
  int rc = set_device(c);
  if (rc) return rc;
  HX_TRY(hipStreamSynchronize(c->stream));
  const int mirror = isx ? 0 : (isy ? 1 : -1);
  if (mirror >= 0 && (rc = build_mirror_view(c, mirror))) return rc;
  if (mirror >= 0 && !(c->sym_stream && c->sym_go && c->sym_done)) {  // system 2's stream (default priority)
    if (!c->sym_stream && hipStreamCreateWithFlags(&c->sym_stream, hipStreamNonBlocking) != hipSuccess)
      c->sym_stream = nullptr;
    if (c->sym_stream && !c->sym_go && hipEventCreateWithFlags(&c->sym_go, hipEventDisableTiming) != hipSuccess)
      c->sym_go = nullptr;
    if (c->sym_stream && !c->sym_done && hipEventCreateWithFlags(&c->sym_done, hipEventDisableTiming) != hipSuccess)
      c->sym_done = nullptr;
    if (!(c->sym_stream && c->sym_go && c->sym_done)) {  // all three or none
      if (c->sym_go) (void)hipEventDestroy(c->sym_go);
      if (c->sym_done) (void)hipEventDestroy(c->sym_done);
      if (c->sym_stream) (void)hipStreamDestroy(c->sym_stream);
      c->sym_stream = nullptr;
      c->sym_go = c->sym_done = nullptr;
      return fail(HAMSX_EDEVICE, "stream/event creation failed on device %d", c->device);
    }
  }
  if (mirror != c->sym_mirror) {
    c->sym_mirror = mirror;
    c->mbins.n = 0;
    c->field_gen++;  // the field rows change layout
    c->have_A = c->have_S = c->have_LU = c->have_X = c->have_C = false;
  }
  return 0;
}

int hamsx_factor(hamsx_ctx* c) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  int rc = set_device(c);
  if (rc) return rc;
  if ((rc = do_factor(c))) return rc;
  int fl[8] = {0};  // INFO in fl[1]; a symmetry plane's system 2 in fl[5]
  HX_TRY(hipMemcpyAsync(fl, c->flags.p, sizeof fl, hipMemcpyDeviceToHost, c->stream));
  HX_TRY(hipStreamSynchronize(c->stream));
  return fl[1] ? fl[1] : (c->sym() ? fl[5] : 0);
}

int hamsx_solve_radiation(hamsx_ctx* c, double* mxpot) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  int rc = set_device(c);
  if (rc || (rc = do_solve(c))) return rc;
  return fetch_mxpot(c, 0, 6 * c->nbody, mxpot);
}

int hamsx_solve_diffraction(hamsx_ctx* c, double* mxpot) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  int rc = set_device(c);
  if (rc || (rc = do_solve(c))) return rc;
  return fetch_mxpot(c, 6 * c->nbody, c->nbeta, mxpot);
}

int hamsx_forces(hamsx_ctx* c, double* amas, double* bdmp, double* exfc) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  int rc = set_device(c);
  if (rc || (rc = do_solve(c))) return rc;
  int info = 0;
  if ((rc = finish_forces(c, amas, bdmp, exfc, &info))) return rc;
  return info;
}

int hamsx_solve_frequency(hamsx_ctx* c, double w1, double h, double* amas, double* bdmp, double* exfc) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  if (!c->have_waves) return fail(HAMSX_ESTATE, "hamsx_solve_frequency needs hamsx_set_waves first");
  double v, wk, w, wvn[HAMSX_NK];
  int rc = hamsx_wave_properties(w1 == 0.0 || w1 == -1.0 ? HAMSX_FREQ_LIMIT : 3, w1, h, &v, &wk, &w, wvn);
  if (rc || (rc = hamsx_set_frequency(c, v, wk, w, h, wvn, HAMSX_NK)) || (rc = hamsx_assemble(c))) return rc;
  if ((rc = do_solve(c))) return rc;
  int info = 0;
  if ((rc = finish_forces(c, amas, bdmp, exfc, &info))) return rc;
  return info;
}

namespace {
// The whole solve after hamsx_set_frequency: assembly, factorisation and
// solves, force reductions and the copies into the staging area.
int enqueue_solve_eager(hamsx_ctx* c) {
  int rc = hamsx_assemble(c);
  if (rc || (rc = do_solve(c)) || (rc = enqueue_forces(c))) return rc;
  return 0;
}

// Recorded launch sequences cover the dense path of one mesh at one depth:
// no lid, no symmetry plane, the LU solver, no stage timing.  The frequency
// reaches the kernels through the device scalars (freq_now), so one
// recording replays for every frequency; what the host decides from the
// frequency (the depth's pair bins, the zero/infinite-frequency branch) and
// where the buffers live form the key.
bool graph_eligible(const hamsx_ctx* c) {
  return c->opt_graphs && c->nlid == 0 && !c->sym() && c->solver == HAMSX_SOLVER_LU && !c->profiling;
}


void note_state(hamsx_ctx* c) {
  c->g_have_A = c->have_A;
  c->g_have_S = c->have_S;
  c->g_have_LU = c->have_LU;
  c->g_have_X = c->have_X;
  c->g_have_C = c->have_C;
}

void restore_state(hamsx_ctx* c) {
  c->have_A = c->g_have_A;
  c->have_S = c->g_have_S;
  c->have_LU = c->g_have_LU;
  c->have_X = c->g_have_X;
  c->have_C = c->g_have_C;
}

// Enqueues the solve: replays the recorded sequence when its key matches;
// the first solve under a new key runs eagerly (it may allocate and sort the
// pair bins); the second is recorded with stream capture and replayed.  A
// recording that fails (or that allocated) is discarded and the solve runs
// eagerly; that key is then not recorded again.
int enqueue_solve(hamsx_ctx* c) {
  if (!graph_eligible(c)) return enqueue_solve_eager(c);
  // the bins of this depth must exist before the key is taken
  const hamsx_ctx::GraphKey k0 = c->graph_key_now();
  if (c->graph_exec && c->graph_key == k0) {
    HX_TRY(hipGraphLaunch(c->graph_exec, c->stream));
    restore_state(c);
    c->graph_launches++;
    return 0;
  }
  if (!c->warm || !(c->warm_key == k0) || c->graph_refused) {
    const bool same = c->warm && c->warm_key == k0;
    int rc = enqueue_solve_eager(c);
    if (rc) return rc;
    if (!same) {
      c->graph_refused = false;
      c->warm = true;
      c->warm_key = c->graph_key_now();  // after the eager solve's allocations and pair sort
    }
    return 0;
  }
  // record
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  // The recording holds the one-stream LU schedule (the same factors bit for
  // bit): replays of a recording that forks the look-ahead stream measured
  // 12 % slower than eager launches with four solves in flight, one-stream
  // recordings as fast as eager.
  HX_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const bool A0 = c->have_A, S0 = c->have_S, LU0 = c->have_LU, X0 = c->have_X, C0 = c->have_C;
  c->capturing = true;
  int rc = enqueue_solve_eager(c);
  c->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  hipGraphExec_t x = nullptr;
  bool ok = rc == 0 && e == hipSuccess && g && c->graph_key_now() == k0;
  if (ok && hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) ok = false;
  if (!ok) {
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    c->graph_refused = true;
    c->have_A = A0;
    c->have_S = S0;
    c->have_LU = LU0;
    c->have_X = X0;
    c->have_C = C0;
    if (rc) return rc;
    return enqueue_solve_eager(c);
  }
  note_state(c);
  c->graph = g;
  c->graph_exec = x;
  c->graph_key = k0;
  c->graph_records++;
  HX_TRY(hipGraphLaunch(x, c->stream));
  c->graph_launches++;
  return 0;
}
}  // namespace

// Completion notices: a host function at the end of each submitted solve
// appends the context's tag; hamsx_queue_wait blocks on them.
struct hamsx_queue {
  std::mutex m;
  std::condition_variable cv;
  std::deque<int> done;
};

namespace {
void notify_done(void* p) {
  auto* c = static_cast<hamsx_ctx*>(p);
  hamsx_queue* q = c->queue;
  {
    std::lock_guard<std::mutex> lk(q->m);
    q->done.push_back(c->queue_tag);
  }
  q->cv.notify_all();
}
}  // namespace

int hamsx_submit_frequency(hamsx_ctx* c, double w1, double h) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  if (!c->have_waves) return fail(HAMSX_ESTATE, "hamsx_submit_frequency needs hamsx_set_waves first");
  double v, wk, w, wvn[HAMSX_NK];
  int rc = hamsx_wave_properties(w1 == 0.0 || w1 == -1.0 ? HAMSX_FREQ_LIMIT : 3, w1, h, &v, &wk, &w, wvn);
  if (rc || (rc = hamsx_set_frequency(c, v, wk, w, h, wvn, HAMSX_NK))) return rc;
  if ((rc = enqueue_solve(c))) return rc;
  if (c->queue) HX_TRY(hipLaunchHostFunc(c->stream, notify_done, c));
  c->pending = true;
  return 0;
}

static_assert(HAMSX_GROUP_MAX == hamsx::kLuBatchMax, "a frequency group is one batched factorisation");

int hamsx_submit_frequency_group(hamsx_ctx** ctx, int count, const double* w1, double h) {
  if (!ctx || !w1) return fail(HAMSX_EARG, "ctx or w1 is NULL");
  if (count < 1 || count > hamsx::kLuBatchMax)
    return fail(HAMSX_EARG, "a frequency group holds 1 to %d contexts", hamsx::kLuBatchMax);
  hamsx_ctx* c0 = ctx[0];
  for (int i = 0; i < count; i++) {
    hamsx_ctx* c = ctx[i];
    if (!c) return fail(HAMSX_EARG, "context %d of the group is NULL", i);
    for (int j = 0; j < i; j++)
      if (ctx[j] == c) return fail(HAMSX_EARG, "context %d appears twice in the group", i);
    if (c->pending) return busy(c);
    if (!c->have_waves) return fail(HAMSX_ESTATE, "hamsx_submit_frequency_group needs hamsx_set_waves first");
    if (c->device != c0->device || c->n != c0->n || c->prec != c0->prec)
      return fail(HAMSX_EARG, "the contexts of a group share one device, one mesh size and one precision");
    if (c->solver != HAMSX_SOLVER_LU || c->sym() || c->nlid > 0)
      return fail(HAMSX_EARG, "a frequency group takes the dense LU path (no GMRES, lid or symmetry plane)");
  }
  const bool c64 = c0->prec == HAMSX_PREC_MIXED;
  const int n = c0->n;
  // each context: its frequency, assembly and right-hand sides on its own stream
  for (int i = 0; i < count; i++) {
    hamsx_ctx* c = ctx[i];
    double v, wk, w, wvn[HAMSX_NK];
    int rc = hamsx_wave_properties(w1[i] == 0.0 || w1[i] == -1.0 ? HAMSX_FREQ_LIMIT : 3, w1[i], h, &v, &wk, &w, wvn);
    if (rc || (rc = hamsx_set_frequency(c, v, wk, w, h, wvn, HAMSX_NK)) || (rc = hamsx_assemble(c)) ||
        (rc = do_rhs(c)) || (rc = factor_buffers(c)))
      return rc;
    if (!c->grp_ev) HX_TRY(hipEventCreateWithFlags(&c->grp_ev, hipEventDisableTiming));
    if (c64) hamsx::launch::to_c64(c->stream, system_matrix(c), c->Af.p, (size_t)n * n);
    if ((rc = check_launch())) return rc;
    HX_TRY(hipEventRecord(c->grp_ev, c->stream));
  }
  // one factorisation launch sequence for the whole group, on the first context's stream
  int rc = set_device(c0);
  if (rc) return rc;
  for (int i = 1; i < count; i++) HX_TRY(hipStreamWaitEvent(c0->stream, ctx[i]->grp_ev, 0));
  hamsx::LuBatch b{};
  for (int i = 0; i < count; i++) {
    hamsx_ctx* c = ctx[i];
    b.m[i] = hamsx::LuMat{c64 ? static_cast<void*>(c->Af.p) : static_cast<void*>(system_matrix(c)),
                          c->ipiv.p, c->perm.p, c->widx.p, c->wcnt.p, c->ludiag.p, c->flags.p + 1};
  }
  const hamsx::LuAux o = lu_options(c0, true, count > 1);
  {
    StageTimer tm(c0, 2);
    hamsx::launch::lu_factor_batch(c0->stream, b, count, n, c64, &o);
  }
  if ((rc = check_launch())) return rc;
  // With one right-hand-side count (and in mixed precision one refinement
  // setting): the group's triangular solves, or its refined complex64 solves,
  // as one batched launch sequence too
  bool batched = true;
  for (int i = 1; i < count && batched; i++)
    batched = ctx[i]->nrhs() == c0->nrhs() &&
              (!c64 || (ctx[i]->refine_iters == c0->refine_iters && ctx[i]->refine_tol == c0->refine_tol));
  if (batched) {
    StageTimer tm(c0, 3);
    if (!c64) {
      hamsx::SolveBatch sb{};
      for (int i = 0; i < count; i++) {
        hamsx_ctx* c = ctx[i];
        sb.m[i] = hamsx::SolveMat{system_matrix(c), c->perm.p, c->X.p, c->Xtmp.p, c->trinv.p};
      }
      hamsx::launch::lu_solve_batch(c0->stream, sb, count, n, c0->nrhs());
    } else {
      hamsx::RefineBatch rb{};
      for (int i = 0; i < count; i++) {
        hamsx_ctx* c = ctx[i];
        rb.m[i] = hamsx::RefineMat{system_matrix(c), c->Af.p, c->perm.p, c->X.p, c->Xtmp.p, c->trinv.p,
                                   c->B.p, c->D.p, c->R.p, c->mvw.p, c->norms.p, c->flags.p + 2};
      }
      hamsx::launch::refine_solve_batch(c0->stream, rb, count, n, c0->nrhs(), c0->refine_iters, c0->refine_tol);
    }
    if ((rc = check_launch())) return rc;
  }
  HX_TRY(hipEventRecord(c0->grp_ev, c0->stream));
  // each context: its solves (unless batched above) and force reductions on its own stream
  for (int i = 0; i < count; i++) {
    hamsx_ctx* c = ctx[i];
    if ((rc = set_device(c))) return rc;
    if (i > 0) HX_TRY(hipStreamWaitEvent(c->stream, c0->grp_ev, 0));
    c->have_LU = true;
    if (!c64) c->have_A = false;  // factored in place
    if (batched) c->have_X = true;
    else if ((rc = solve_factored(c))) return rc;
    if ((rc = enqueue_forces(c))) return rc;
    if (c->queue) HX_TRY(hipLaunchHostFunc(c->stream, notify_done, c));
    c->pending = true;
  }
  return 0;
}

int hamsx_queue_create(hamsx_queue** q) {
  if (!q) return fail(HAMSX_EARG, "queue pointer is NULL");
  *q = new hamsx_queue();
  return 0;
}

int hamsx_queue_destroy(hamsx_queue* q) {
  delete q;
  return 0;
}

int hamsx_set_queue(hamsx_ctx* c, hamsx_queue* q, int tag) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  c->queue = q;
  c->queue_tag = tag;
  return 0;
}

int hamsx_queue_wait(hamsx_queue* q, double timeout_s, int* tag) {
  if (!q || !tag) return fail(HAMSX_EARG, "queue or tag is NULL");
  std::unique_lock<std::mutex> lk(q->m);
  const auto pred = [q] { return !q->done.empty(); };
  if (timeout_s < 0.0) q->cv.wait(lk, pred);
  else if (!q->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return 0;
  *tag = q->done.front();
  q->done.pop_front();
  return 1;
}

int hamsx_graph_stats(hamsx_ctx* c, unsigned long long* records, unsigned long long* launches) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (records) *records = c->graph_records;
  if (launches) *launches = c->graph_launches;
  return 0;
}

int hamsx_query(hamsx_ctx* c) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (!c->pending) return 1;
  int rc = set_device(c);
  if (rc) return rc;
  hipError_t e = hipStreamQuery(c->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return fail(HAMSX_EDEVICE, "hipStreamQuery: %s", hipGetErrorString(e));
}

int hamsx_collect(hamsx_ctx* c, double* amas, double* bdmp, double* exfc) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (!c->pending) return fail(HAMSX_ESTATE, "nothing submitted on this context");
  int rc = set_device(c);
  if (rc) return rc;
  c->pending = false;
  int info = 0;
  if ((rc = complete_forces(c, amas, bdmp, exfc, &info))) return rc;
  return info;
}

int hamsx_set_field_points(hamsx_ctx* c, int nfp, const double* xfp) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  if (c->n == 0) return fail(HAMSX_ESTATE, "hamsx_set_field_points needs hamsx_set_mesh first");
  if (nfp < 0 || (nfp > 0 && !xfp)) return fail(HAMSX_EARG, "bad field-point arguments");
  for (size_t k = 0; k < 3 * (size_t)nfp; k++)
    if (!std::isfinite(xfp[k])) return fail(HAMSX_EARG, "field point %zu is not finite", k % nfp + 1);
  int rc = set_device(c);
  if (rc) return rc;
  HX_TRY(hipStreamSynchronize(c->stream));
  if (nfp > 0) {
    if ((rc = c->fpx.ensure(nfp)) || (rc = c->fpy.ensure(nfp)) || (rc = c->fpz.ensure(nfp))) return rc;
    HX_TRY(hipMemcpy(c->fpx.p, xfp, sizeof(double) * nfp, hipMemcpyHostToDevice));
    HX_TRY(hipMemcpy(c->fpy.p, xfp + nfp, sizeof(double) * nfp, hipMemcpyHostToDevice));
    HX_TRY(hipMemcpy(c->fpz.p, xfp + 2 * (size_t)nfp, sizeof(double) * nfp, hipMemcpyHostToDevice));
  }
  c->nfp = nfp;
  c->field_gen++;
  return 0;
}

int hamsx_field_values(hamsx_ctx* c, int kind, double* values, double* incident, int* onbody) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  if (kind < HAMSX_FIELD_POTENTIAL || kind > HAMSX_FIELD_AUTO) return fail(HAMSX_EARG, "unknown field kind %d", kind);
  if (c->nfp == 0) return fail(HAMSX_ESTATE, "no field points: call hamsx_set_field_points first");
  int rc = set_device(c);
  if (rc || (rc = do_solve(c))) return rc;
  const int n = c->n, nfp = c->nfp, nd = 6 * c->nbody, nr = c->nrhs(), nbeta = c->nbeta;
  const bool finite = c->freq.h >= 0.0;
  const bool binned = finite && n < 0xFFFF && !force_grid_assembly();
  // with a symmetry plane every point brings its mirror image: images = 2
  // rows per point, and the mirror rows meet the second column set Z2
  const int images = c->sym() ? 2 : 1;
  const int F = std::max(1, field_chunk(c, binned) / images);
  const bool want_inc = incident && nbeta > 0;
  const size_t dn = (size_t)n * std::max(1, nbeta);
  if ((rc = c->fDincp.ensure(dn * images)) || (rc = c->fZ.ensure((size_t)2 * n * nr)) ||
      (c->sym() && (rc = c->fZ2.ensure((size_t)2 * n * nr))) ||
      (rc = c->fM.ensure((size_t)2 * n * images * F)) || (rc = c->fP.ensure((size_t)F * nr)) ||
      (rc = c->fOut.ensure((size_t)nfp * nr)) || (rc = c->fOnb.ensure(nfp)) ||
      (rc = c->ffx.ensure(n + images * F)) || (rc = c->ffy.ensure(n + images * F)) ||
      (rc = c->ffz.ensure(n + images * F)) || (want_inc && (rc = c->fInc.ensure((size_t)nfp * nbeta))) ||
      (finite && (rc = c->fmodetab.ensure(hamsx::mode_tab_len(n + images * F)))) || (rc = blas_handle(c)))
    return rc;
  const hamsx::MeshDev m = c->mesh();
  StageTimer tm(c, 9);
  hamsx::launch::dincp(c->stream, m, c->freq, c->waves(), c->fDincp.p);
  hamsx::launch::field_weights(c->stream, m, c->nbody, nbeta, c->fDincp.p, c->X.p, c->fZ.p);
  if (c->sym()) {
    // antisymmetric modes: x = 0: surge, pitch, yaw; y = 0: sway, roll, yaw
    const int flip = c->sym_mirror == 0 ? 0x31 : 0x2A;
    hamsx::launch::dincp(c->stream, m, c->freq, c->waves(), c->fDincp.p + dn, c->sym_mirror);
    hamsx::launch::field_weights(c->stream, m, c->nbody, nbeta, c->fDincp.p + dn, c->X2.p, c->fZ2.p, flip);
  }
  if ((rc = check_launch())) return rc;
  HX_TRY(hipMemcpyAsync(c->ffx.p, c->cx.p, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  HX_TRY(hipMemcpyAsync(c->ffy.p, c->cy.p, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  HX_TRY(hipMemcpyAsync(c->ffz.p, c->cz.p, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  hamsx::FreqDev ff = c->freq;
  ff.tab = finite ? c->fmodetab.p : nullptr;
  int* ovf = c->flags.p + 4;
  const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
  for (int p0 = 0, chunk = 0; p0 < nfp; p0 += F, chunk++) {
    const int nc = std::min(F, nfp - p0);
    HX_TRY(hipMemcpyAsync(c->ffx.p + n, c->fpx.p + p0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
    HX_TRY(hipMemcpyAsync(c->ffy.p + n, c->fpy.p + p0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
    HX_TRY(hipMemcpyAsync(c->ffz.p + n, c->fpz.p + p0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
    if (c->sym())
      hamsx::launch::reflect_points(c->stream, nc, c->sym_mirror, c->fpx.p + p0, c->fpy.p + p0, c->fpz.p + p0,
                                    c->ffx.p + n + nc, c->ffy.p + n + nc, c->ffz.p + n + nc);
    // the field view: sources are the hull panels, rows n .. n + nc - 1 the
    // chunk's points (then their mirror images); FLAG from the point itself
    const int rows = images * nc;
    hamsx::MeshDev fm = m;
    fm.fx = fm.gx = c->ffx.p;
    fm.fy = fm.gy = c->ffy.p;
    fm.fz = fm.gz = c->ffz.p;
    fm.nf = n + rows;
    fm.row0 = n;
    double2* S = c->fM.p;
    double2* T = c->fM.p + (size_t)rows * n;
    if (finite && c->freq.v != 0.0 && c->freq.v != -1.0) hamsx::launch::mode_tables(c->stream, fm, ff);
    if (binned) {
      if ((rc = ensure_bins(c, c->fbins, fm, (c->field_gen << 20) | (unsigned long long)chunk))) return rc;
      assemble_binned(c, c->fbins, fm, ff, T, S, ovf, false);
    } else {
      hamsx::launch::assemble(c->stream, fm, ff, T, S, ovf);
    }
    hamsx::launch::on_body(c->stream, m, c->vtx.p, nc, c->fpx.p + p0, c->fpy.p + p0, c->fpz.p + p0, c->fOnb.p + p0);
    if (c->sym())  // the free term counts the mirror image too
      hamsx::launch::on_body(c->stream, m, c->vtx.p, nc, c->ffx.p + n + nc, c->ffy.p + n + nc, c->ffz.p + n + nc,
                             c->fOnb.p + p0, true);
    if ((rc = check_launch())) return rc;
    // P = M1 Z1 (+ M2 Z2): M = [S | T] with leading dimension rows, M2 its mirror rows
    for (int im = 0; im < images; im++)
      if (rocblas_zgemm(c->blas, rocblas_operation_none, rocblas_operation_none, nc, nr, 2 * n, &one,
                        reinterpret_cast<const rocblas_double_complex*>(c->fM.p + (size_t)im * nc), rows,
                        reinterpret_cast<const rocblas_double_complex*>(im == 0 ? c->fZ.p : c->fZ2.p), 2 * n,
                        im == 0 ? &zero : &one, reinterpret_cast<rocblas_double_complex*>(c->fP.p),
                        nc) != rocblas_status_success)
        return fail(HAMSX_EDEVICE, "rocblas_zgemm (field points) failed");
    hamsx::launch::field_finish(c->stream, nc, c->fpx.p + p0, c->fpy.p + p0, c->fpz.p + p0, c->fOnb.p + p0, c->freq,
                                c->waves(), nd, kind, c->fP.p, (size_t)nfp, c->fOut.p + p0,
                                want_inc ? c->fInc.p + p0 : nullptr);
    if ((rc = check_launch())) return rc;
    int ov = 0;
    HX_TRY(hipMemcpyAsync(&ov, ovf, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HX_TRY(hipStreamSynchronize(c->stream));
    if (ov & 1) return fail(HAMSX_EDEVICE, "adaptive quadrature stack exhausted in the finite-depth kernel");
  }
  if (values)
    HX_TRY(hipMemcpyAsync(values, c->fOut.p, sizeof(double2) * (size_t)nfp * nr, hipMemcpyDeviceToHost, c->stream));
  if (want_inc)
    HX_TRY(hipMemcpyAsync(incident, c->fInc.p, sizeof(double2) * (size_t)nfp * nbeta, hipMemcpyDeviceToHost,
                          c->stream));
  if (onbody) HX_TRY(hipMemcpyAsync(onbody, c->fOnb.p, sizeof(int) * nfp, hipMemcpyDeviceToHost, c->stream));
  HX_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int hamsx_set_body_matrices(hamsx_ctx* c, const double* mass, const double* blnr, const double* bqdr,
                            const double* crs, const double* kstf) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  if (c->n == 0) return fail(HAMSX_ESTATE, "hamsx_set_body_matrices needs hamsx_set_mesh first");
  if (!mass || !blnr || !bqdr || !crs || !kstf) return fail(HAMSX_EARG, "body matrices must not be NULL");
  const int nb = c->nbody, nd = 6 * nb;
  if (nd > 192) return fail(HAMSX_EARG, "the motion solve supports at most 32 bodies");
  int rc = set_device(c);
  if (rc) return rc;
  HX_TRY(hipStreamSynchronize(c->stream));
  const size_t nn = (size_t)nd * nd;
  std::vector<double> h(5 * nn, 0.0);
  const double* src[5] = {mass, blnr, bqdr, crs, kstf};
  double qmax = 0.0;
  // LOCALTOGLOBAL (SolveMotionMulti.f90:24-40): body b's 6 x 6 block on the diagonal
  for (int q = 0; q < 5; q++)
    for (int b = 0; b < nb; b++)
      for (int j = 0; j < 6; j++)
        for (int i = 0; i < 6; i++) {
          const double v = src[q][36 * (size_t)b + i + 6 * j];
          if (!std::isfinite(v)) return fail(HAMSX_EARG, "body matrix entry is not finite");
          h[q * nn + (size_t)(6 * b + i) + (size_t)nd * (6 * b + j)] = v;
          if (q == 2) qmax = std::max(qmax, std::fabs(v));
        }
  if ((rc = c->bmat.ensure(5 * nn))) return rc;
  HX_TRY(hipMemcpy(c->bmat.p, h.data(), sizeof(double) * 5 * nn, hipMemcpyHostToDevice));
  c->bodymat_nbody = nb;
  c->motion_quad = qmax >= 1.0e-6 ? 1 : 0;  // DLANGE('M', BQDR) >= 1e-6 (:80, :89)
  return 0;
}

int hamsx_motions(hamsx_ctx* c, double* dspl, int* iters) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  if (c->bodymat_nbody == 0 || c->bodymat_nbody != c->nbody)
    return fail(HAMSX_ESTATE, "hamsx_motions needs hamsx_set_body_matrices for the current mesh");
  int rc = set_device(c);
  if (rc || (rc = do_solve(c))) return rc;
  const int nd = 6 * c->nbody, nbeta = c->nbeta;
  if (nbeta == 0) return 0;
  if ((rc = c->mwork.ensure((size_t)nd * nd * nbeta)) || (rc = c->mdspl.ensure((size_t)nd * nbeta)) ||
      (rc = c->miters.ensure(nbeta)))
    return rc;
  hamsx::launch::forces(c->stream, c->mesh(), c->freq, c->waves(), c->nbody, c->body_start_d.p, c->X.p, c->out.p,
                        c->sym() ? c->X2.p : nullptr, c->sym_mirror);
  hamsx::launch::motions(c->stream, nd, nbeta, c->freq.w1, c->out.p, c->bmat.p, c->motion_quad, c->mwork.p,
                         c->mdspl.p, c->miters.p);
  if ((rc = check_launch())) return rc;
  std::vector<int> it(nbeta);
  if (dspl)
    HX_TRY(hipMemcpyAsync(dspl, c->mdspl.p, sizeof(double2) * (size_t)nd * nbeta, hipMemcpyDeviceToHost, c->stream));
  HX_TRY(hipMemcpyAsync(it.data(), c->miters.p, sizeof(int) * nbeta, hipMemcpyDeviceToHost, c->stream));
  HX_TRY(hipStreamSynchronize(c->stream));
  if (iters) std::copy(it.begin(), it.end(), iters);
  for (int b = 0; b < nbeta; b++)
    if (it[b] < 0) return fail(HAMSX_ENOCONV, "quadratic-damping iteration of heading %d did not converge", b + 1);
  return 0;
}

int hamsx_get_matrix(hamsx_ctx* c, int which, double* out, int* ipiv) {
  if (!c || !out) return fail(HAMSX_EARG, "bad arguments");
  if (c->pending) return busy(c);
  int rc = set_device(c);
  if (rc) return rc;
  const size_t nn = (size_t)c->n * c->n;
  if (which == 0 || which == 2) {
    if (which == 0 && !c->have_A) return fail(HAMSX_ESTATE, "A is not available (not assembled, or factored in place)");
    if (which == 0 && c->nlid > 0) {  // with a lid: C = A^T A, the matrix that is factored
      if (c->have_LU && c->prec == HAMSX_PREC_FP64)
        return fail(HAMSX_ESTATE, "C is factored in place; read it before hamsx_factor");
      int rc2 = ensure_normal_matrix(c);
      if (rc2) return rc2;
      HX_TRY(hipMemcpyAsync(out, c->Cm.p, sizeof(double2) * nn, hipMemcpyDeviceToHost, c->stream));
      HX_TRY(hipStreamSynchronize(c->stream));
      return 0;
    }
    if (which == 2 && !c->have_LU) return fail(HAMSX_ESTATE, "no factorisation available");
    if (which == 2 && c->prec == HAMSX_PREC_MIXED) {
      std::vector<float2> tmp(nn);
      HX_TRY(hipMemcpyAsync(tmp.data(), c->Af.p, sizeof(float2) * nn, hipMemcpyDeviceToHost, c->stream));
      HX_TRY(hipStreamSynchronize(c->stream));
      for (size_t k = 0; k < nn; k++) {
        out[2 * k] = tmp[k].x;
        out[2 * k + 1] = tmp[k].y;
      }
    } else {
      HX_TRY(hipMemcpyAsync(out, c->A.p, sizeof(double2) * nn, hipMemcpyDeviceToHost, c->stream));
    }
    if (which == 2 && ipiv) HX_TRY(hipMemcpyAsync(ipiv, c->ipiv.p, sizeof(int) * c->n, hipMemcpyDeviceToHost, c->stream));
  } else if (which == 1) {
    if (!c->have_S) return fail(HAMSX_ESTATE, "S is not available");
    HX_TRY(hipMemcpyAsync(out, c->S.p, sizeof(double2) * nn, hipMemcpyDeviceToHost, c->stream));
  } else {
    return fail(HAMSX_EARG, "which must be 0 (A), 1 (S) or 2 (LU)");
  }
  HX_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int hamsx_set_solver(hamsx_ctx* c, int solver, double tol, int restart, int maxiter) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  drop_graph(c);
  if (solver != HAMSX_SOLVER_LU && solver != HAMSX_SOLVER_GMRES) return fail(HAMSX_EARG, "unknown solver %d", solver);
  if (solver == HAMSX_SOLVER_GMRES) {
    if (!(tol > 0.0) || restart < 1 || restart > 200 || maxiter < 1)
      return fail(HAMSX_EARG, "GMRES needs tol > 0, 1 <= restart <= 200, maxiter >= 1");
    c->gm_tol = tol;
    c->gm_restart = restart;
    c->gm_maxiter = maxiter;
  }
  c->solver = solver;
  c->have_X = false;
  return 0;
}

int hamsx_set_refinement(hamsx_ctx* c, int max_steps, double tol) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  drop_graph(c);
  if (max_steps < 0 || !(tol >= 0.0)) return fail(HAMSX_EARG, "max_steps must be >= 0 and tol >= 0");
  c->refine_iters = max_steps;
  c->refine_tol = tol;
  c->have_X = false;
  return 0;
}

int hamsx_set_hmatrix(hamsx_ctx* c, int enable, double eta, double aca_tol, int leaf, int rank_max) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (c->pending) return busy(c);
  drop_graph(c);
  if (!(eta > 0.0) || !(aca_tol > 0.0) || leaf < 2 || rank_max < 1 || rank_max > 64)
    return fail(HAMSX_EARG, "H-matrix: need eta > 0, aca_tol > 0, leaf >= 2, 1 <= rank_max <= 64");
  c->hm_on = enable != 0;
  c->hm_eta = eta;
  c->hm_tol = aca_tol;
  c->hm_leaf = leaf;
  c->hm_rank = rank_max;
  c->hm_part = c->hm_built = false;
  c->have_X = false;
  return 0;
}

int hamsx_hmatrix_stats(hamsx_ctx* c, int* nlowrank, int* ndense, double* mean_rank, double* compression) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (nlowrank) *nlowrank = c->hm_built ? c->hm_nlr : 0;
  if (ndense) *ndense = c->hm_built ? c->hm_ndense : 0;
  if (mean_rank) *mean_rank = c->hm_built ? c->hm_mean_rank : 0.0;
  if (compression) *compression = c->hm_built ? c->hm_ratio : 0.0;
  return 0;
}

int hamsx_hmatrix_matvec(hamsx_ctx* c, const double* x, int nrhs, double* y) {
  if (!c || !x || !y || nrhs <= 0) return fail(HAMSX_EARG, "hamsx_hmatrix_matvec: bad arguments");
  if (c->pending) return busy(c);
  if (!c->have_A) return fail(HAMSX_ESTATE, "hamsx_hmatrix_matvec needs an assembled, unfactored matrix");
  int rc = set_device(c);
  if (rc || (rc = ensure_normal_matrix(c))) return rc;
  const double2* Aop = system_matrix(c);
  if (!c->hm_built || c->hm_nrhs != nrhs) {
    c->hm_built = false;
    if ((rc = hm_build(c, Aop, nrhs))) return rc;
  }
  const size_t cnt = (size_t)c->n * nrhs;
  if ((rc = c->gU.ensure(cnt)) || (rc = c->gW.ensure(cnt))) return rc;
  HX_TRY(hipMemcpyAsync(c->gU.p, x, sizeof(double2) * cnt, hipMemcpyHostToDevice, c->stream));
  hamsx::launch::hm_matvec(c->stream, hm_dev(c), c->n, c->gU.p, nrhs, nullptr, 1.0, 0.0, c->gW.p, c->hmXp.p,
                           c->hmYp.p);
  HX_TRY(hipMemcpyAsync(y, c->gW.p, sizeof(double2) * cnt, hipMemcpyDeviceToHost, c->stream));
  HX_TRY(hipStreamSynchronize(c->stream));
  return check_launch();
}

int hamsx_hmatrix_partition(int n, const double* cen, int leaf, double eta, int* perm, int max_blocks, int* blocks,
                            int* nblocks) {
  if (n <= 0 || !cen || !perm || !nblocks || leaf < 2 || !(eta > 0.0) || max_blocks < 0 ||
      (max_blocks > 0 && !blocks))
    return fail(HAMSX_EARG, "hamsx_hmatrix_partition: bad arguments");
  std::vector<int> pv;
  std::vector<hamsx::HNode> tree;
  std::vector<hamsx::HBlk> bl;
  hamsx::hmatrix_partition(n, cen, leaf, eta, pv, tree, bl);
  std::copy(pv.begin(), pv.end(), perm);
  *nblocks = (int)bl.size();
  if ((int)bl.size() > max_blocks)
    return fail(HAMSX_EARG, "hamsx_hmatrix_partition: %zu blocks, room for %d", bl.size(), max_blocks);
  for (size_t k = 0; k < bl.size(); k++) {
    int* o = blocks + 5 * k;
    o[0] = bl[k].rs;
    o[1] = bl[k].nr;
    o[2] = bl[k].cs;
    o[3] = bl[k].nc;
    o[4] = bl[k].aca;
  }
  return 0;
}

int hamsx_refine_stats(hamsx_ctx* c, int* iters, double* resid) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  if (iters) *iters = c->last_iters;
  if (resid) *resid = c->last_resid;
  return 0;
}

int hamsx_set_profiling(hamsx_ctx* c, int on) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  drop_graph(c);
  c->profiling = on != 0;
  return 0;
}

int hamsx_stage_times(hamsx_ctx* c, int reset, double* ms, int* count) {
  if (!c) return fail(HAMSX_EARG, "ctx is NULL");
  int rc = set_device(c);
  if (rc) return rc;
  HX_TRY(hipStreamSynchronize(c->stream));
  harvest_marks(c);
  for (int k = 0; k < HAMSX_NSTAGE; k++) {
    if (ms) ms[k] = c->stage_ms[k];
    if (count) count[k] = c->stage_cnt[k];
    if (reset) {
      c->stage_ms[k] = 0.0;
      c->stage_cnt[k] = 0;
    }
  }
  return 0;
}

}  // extern "C"
