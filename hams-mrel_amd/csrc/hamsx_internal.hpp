// hamsx_internal.hpp -- shared declarations between the C-ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace hamsx {

// Launch checks: with HAMSX_DEBUG set in the environment every launch is
// followed by hipGetLastError, and a failure is reported with the kernel name.
bool debug_launches();
void report_launch(const char* name, hipError_t e);
// Records msg as the calling thread's hamsx_last_error and returns code.
int set_error(int code, const char* msg);
#define HX_LAUNCH(kernel, ...)                                                      \
  do {                                                                              \
    hipLaunchKernelGGL(kernel, __VA_ARGS__);                                        \
    if (::hamsx::debug_launches()) ::hamsx::report_launch(#kernel, hipGetLastError()); \
  } while (0)

constexpr int kNK = 200;      // dispersion roots kept (WaveDyn_mod NK)
constexpr int kPanelNB = 32;  // LU panel width
constexpr int kSelRows = 256; // candidate rows per tournament block (one per thread)

// Device-resident, frequency-invariant mesh (structure of arrays).
struct MeshDev {
  int n;                // hull panels (sources; field points 0..n-1)
  int nf;               // field points: the n hull centroids, then the lid centroids (irregular-frequency removal)
  const double *fx, *fy, *fz;  // field-point coordinates, nf each
  const double* cx;
  const double* cy;
  const double* cz;
  const double* area;
  const double* pnsz;
  const double* nrm;    // nrm[m * n + j], m = 0..5
  const int* ncn;
  const double* frame;  // Newman local frame, kFrameLen doubles per panel
  const int* body;      // body index of each panel
  int row0;             // first field row assembled: rows i >= row0 are stored at i - row0, leading
                        // dimension nf - row0 (0 for the system; n for field-point rows)
  const double *gx, *gy, *gz;  // the points FLAG is decided from (nf each): the field points, except
                               // for mirrored rows, which keep the unmirrored centroid
};

// Regions of the finite-depth Green function by R/h (FinGreen3D.f90:183-191)
// and the keys the binned assembly sorts pairs by.  subkey ascends by
// region: LINTON 0; PIDCOCK 1 + (99 - NT) for NT in 49..99; EIGENE
// 52 + (50 - NT) for NT in 10..50; EIGEN 93.
//   * single pairs (field i, source j), one orientation per list entry:
//     key = 2 * subkey + FLAG -- every LINTON pair and every pair whose
//     field point is a lid point;
//   * hull pairs {a < b} outside LINTON, both orientations per list entry:
//     key = kPairKey0 + 4 * (subkey - 1) + 2 * FLAG(a, b) + FLAG(b, a),
//     FLAG(i, j) being the flag of field i and source j.
// Keys depend on the geometry and h only.
enum FinRegion { kRegAny = 0, kRegLinton = 1, kRegPidcock = 2, kRegEigene = 3, kRegEigen = 4 };
constexpr int kSubPidcock = 1, kSubEigene = 52, kSubEigen = 93;
constexpr int kPairKey0 = 2 * kSubEigen + 2;
constexpr int kFinBins = kPairKey0 + 4 * kSubEigen;  // 560
inline constexpr int fin_sub_region(int sub) {
  return sub < kSubPidcock ? kRegLinton : sub < kSubEigene ? kRegPidcock : sub < kSubEigen ? kRegEigene : kRegEigen;
}
inline constexpr bool fin_key_paired(int key) { return key >= kPairKey0; }
inline constexpr int fin_key_region(int key) {
  return fin_sub_region(key < kPairKey0 ? key / 2 : kSubPidcock + (key - kPairKey0) / 4);
}

struct FreqDev {
  double v, wk, w1, h;
  const double* wvn;  // kNK roots on the device
  double* tab;        // finite depth: per-frequency mode tables (layout below)
  // device copy of (v, wk, w1, h) written with the roots: every kernel takes
  // the scalars from here (freq_now), so a launch sequence recorded once
  // serves any frequency of the same depth
  const double* sc = nullptr;
  int far32 = 0;  // EIGEN region's evanescent series in FP32 (mixed precision)
  int limes_ieee = 0;  // diagnostic: every epsilon-table quotient by IEEE division (no fast pass)
};
__host__ __device__ inline FreqDev freq_now(FreqDev f) {
  if (f.sc) {
    f.v = f.sc[0];
    f.wk = f.sc[1];
    f.w1 = f.sc[2];
    f.h = f.sc[3];
  }
  return f;
}

// Per-frequency mode tables of the finite-depth series, kModes modes:
// [pm | ac | ut] (kModes each), nm0 (1), padding to kTabHead, then
// [cz | sz | ct | st] (kModes x n each) and [chz | shz] (n each).
constexpr int kModes = 100;
constexpr size_t kTabHead = 3 * kModes + 8;
inline size_t mode_tab_len(int n) { return kTabHead + (size_t)4 * kModes * n + 2 * (size_t)n; }

struct WaveDev {
  int nbeta;
  const double* beta;  // radians, device
  double amp, xw0, xw1, rho;
  int isol;
};

// H-matrix operator on the device (hmatrix.hip): low-rank blocks
// (ldesc/rank/uoff/voff/U/VH, T = per-block VH x scratch at toff), dense
// blocks (ddesc/doff/D), and per leaf cluster of rows [leaves[k].x,
// leaves[k].y) the blocks covering it, litem[lptr[k] .. lptr[k+1]) (item >= 0
// dense block, ~item low-rank block).  perm: cluster position -> panel.
// columns of a low-rank block per work item of the T = VH x pass
constexpr int kVxCols = 256;
// blocks of a leaf's list per work item of the per-leaf pass
constexpr int kLeafSeg = 8;
struct HmDev {
  const int* perm;
  int nlr, nleaf;
  // T = VH x work items of the low-rank blocks with rank > 0, grouped by
  // rank (<= 8, 16, 32, 64): group g is items vx_start[g] .. vx_start[g+1];
  // item k covers columns [chunk kVxCols, (chunk + 1) kVxCols) of block
  // vx_item[k] = (block, chunk) and writes its rank x nrhs product at
  // T + vx_out[k]: the block's T itself when it has one chunk, else a partial
  // that k_hm_vx_reduce adds up, chunk by chunk, for the vx_nred blocks
  // vx_red[r] = (block, chunks) with partials from T + vx_rbase[r]
  const int2* vx_item;
  const size_t* vx_out;
  int vx_start[5];
  const int2* vx_red;
  const size_t* vx_rbase;
  int vx_nred;
  const int4* ldesc;
  const int* rank;
  const size_t *uoff, *voff, *toff;
  const double2 *U, *VH;
  double2* T;
  const int4* ddesc;
  const size_t* doff;
  const double2* D;
  const int2* leaves;
  const int *lptr, *litem;
  // the per-leaf pass in segments of at most kLeafSeg blocks of a leaf's
  // list: seg[k] = (leaf, first item, end item, partial?) -- a leaf of one
  // segment writes Yp itself, the others write rows x nrhs partials at
  // Ypart + seg_out[k], which k_hm_leaf_reduce adds up for the leaf_nred
  // leaves leaf_red[r] = (leaf, segments) from Ypart + leaf_rbase[r]
  const int4* seg;
  const size_t* seg_out;
  int nseg;
  double2* Ypart;
  const int2* leaf_red;
  const size_t* leaf_rbase;
  int leaf_nred;
  // complex64 copies of U, VH and D (mixed precision), or null
  const float2 *U32 = nullptr, *VH32 = nullptr, *D32 = nullptr;
};

// Look-ahead resources of the LU: a second stream (default priority) and two
// events; null or s2 == null factors on the one stream.
// LU schedule and shape options.  s2 (with its two events) set: look-ahead
// on a second stream.  sel_rows: rows per pivot-selection block (256, 512 or
// 1024).  outer_min_n: outer blocks of four panels from this order on (<= 0:
// never).
struct LuAux {
  hipStream_t s2 = nullptr;
  hipEvent_t panel = nullptr, update = nullptr;
  int sel_rows = kSelRows;
  int outer_min_n = 6144;
};
// Outer blocks from this order on in a batched factorisation of two or more
// matrices (a frequency group), unless the option says otherwise: with the
// pivot selection's latency shared by the batch, the rank-128 far updates
// pay from small orders on (N = 3,968, sixteen matrices: 6.2 -> 5.2 ms per
// matrix in FP64).
constexpr int kOuterMinNGroup = 1024;

// One matrix of a batched factorisation: the n x n matrix (complex128 or
// complex64), its pivots, permutation, tournament workspace (2 nsets0 kPanelNB
// + 2 nsets0 ints, nsets0 = ceil(n / kSelRows)), diagonal-block scratch and
// INFO.  Every matrix of a batch has the same order, so one launch sequence
// factors them all (grid.y = matrix); the descriptors travel in the kernel
// arguments, so a batch needs no device memory of its own.
constexpr int kLuBatchMax = 16;
struct LuMat {
  void* A;
  int* ipiv;
  int* perm;
  int* widx;
  int* wcnt;
  void* diag;
  int* info;
};
struct LuBatch {
  LuMat m[kLuBatchMax];
};
// one factored complex128 system of a batched solve: factors, row
// permutation, right-hand sides (overwritten by the solution), a work copy of
// them and the inverted diagonal blocks (tri_inverse_len(n) entries)
struct SolveMat {
  const double2* LU;
  const int* perm;
  double2* X;
  double2* Xtmp;
  double2* inv;
};
struct SolveBatch {
  SolveMat m[kLuBatchMax];
};
// one system of a batched refined complex64 solve: the FP64 matrix, its
// complex64 factors and permutation, the iterate X (right-hand sides in,
// solution out), the work vectors of the refinement (Xtmp, B, D, R), the
// inverted diagonal blocks, the matvec workspace P, the residual norms and
// the refinement state (converged flag, correction count)
struct RefineMat {
  const double2* M;
  const float2* LU;
  const int* perm;
  double2* X;
  double2* Xtmp;
  double2* inv;
  double2* B;
  double2* D;
  double2* R;
  double2* P;
  double* norms;
  int* state;
};
struct RefineBatch {
  RefineMat m[kLuBatchMax];
};

// GMRES(m) Hessenberg / Givens state on the device (krylov.hip)
struct GmDev {
  double2 *H, *g, *sn;  // (m + 1) m, (m + 1), m per right-hand side
  double* cs;           // m per right-hand side
  double* bnorm;
  int *act, *done, *mact;
  int* iters;           // [0]: iterations over the restarts (longest cycle each)
};

namespace launch {
// influence matrices A and S (complex128, column-major, leading dimension n)
void assemble(hipStream_t s, const MeshDev& m, const FreqDev& f, double2* A, double2* S, int* flags);
// finite depth: fill f.tab for the current frequency (before either assembly)
void mode_tables(hipStream_t s, const MeshDev& m, const FreqDev& f);
// finite depth, binned: count[kFinBins] += pairs per key; then scatter the
// packed (j << 16 | i) pair indices to list[cursor[key]++]
void bin_count(hipStream_t s, const MeshDev& m, double h, int* count);
void bin_scatter(hipStream_t s, const MeshDev& m, double h, int* cursor, unsigned* list);
// one region's kernel (1 LINTON, 2 PIDCOCK, 3 EIGENE, 4 EIGEN) over its slice
// of the sorted pair list: single pairs, or (paired) both orientations of
// hull pairs
void assemble_region(hipStream_t s, const MeshDev& m, const FreqDev& f, int region, bool paired,
                     const unsigned* list, size_t count, double2* A, double2* S, int* flags);
void clear_flags(hipStream_t s, int* flags);
// right-hand sides X(n, 6*nbody + nbeta) = S * generator columns
// (rows: the m.nf field points; S and X have leading dimension m.nf)
void rhs(hipStream_t s, const MeshDev& m, const FreqDev& f, const WaveDev& w, int nbody, const int* body_start_h,
         const int* body_start_d, const double2* S, double2* dum2, double2* partial, int nchunk,
         const int* chunk_body_d, const int* chunk_j0_d, const int* chunk_j1_d, double2* X);
// LU with tournament pivoting; ipiv 1-based, perm = row permutation, info[0] = first zero pivot;
// diag: kPanelNB x kPanelNB complex128 scratch for the factored diagonal block.  With aux, each
// step updates the next panel's columns first and the rest of the trailing matrix on aux->s2,
// overlapping the next pivot selection; the factors are the same bit for bit.
void lu_factor(hipStream_t s, double2* A, int n, int* ipiv, int* perm, int* work_idx, int* work_cnt, int* info,
               double2* diag, const LuAux* aux = nullptr);
// X <- A^{-1} X using the factors, nrhs columns (leading dimension n);
// inv: tri_inverse_len(n) complex128 of workspace (the inverted diagonal blocks)
size_t tri_inverse_len(int n);
void lu_solve(hipStream_t s, const double2* LU, int n, const int* perm, double2* X, double2* Xtmp, double2* inv,
              int nrhs);
// lu_solve for count <= kLuBatchMax factored systems of order n at once
// (each launch covers all of them, grid.z = system); each X is what lu_solve
// gives it alone, bit for bit.
void lu_solve_batch(hipStream_t s, const SolveBatch& b, int count, int n, int nrhs);
// AMAS/BDMP/EXFC reductions into out[] (see hamsx_api.cpp for the layout)
// Xm / mirror: the mirror half's potentials and the symmetry plane
// (0: x = 0, 1: y = 0); mirror = -1 without symmetry.
void forces(hipStream_t s, const MeshDev& m, const FreqDev& f, const WaveDev& w, int nbody, const int* body_start_d,
            const double2* X, double* out, const double2* Xm = nullptr, int mirror = -1);
// Geometric symmetry: A, T2 <- A + T2, A - T2; S, S2 <- S + S2, S - S2 (count entries each)
void sym_combine(hipStream_t s, double2* A, double2* T2, double2* S, double2* S2, size_t count);
// Right-hand-side generators of the two symmetry systems into W1 / W2 (n x (6 nbody + nbeta));
// dinc: scratch of 2 n nbeta
void sym_weights(hipStream_t s, const MeshDev& m, const FreqDev& f, const WaveDev& w, int nbody, int mirror,
                 double2* dinc, double2* W1, double2* W2);
// X, X2 <- (X + X2) / 2, (X - X2) / 2
void sym_recombine(hipStream_t s, double2* X, double2* X2, size_t count);
// single-precision copy of A for the mixed path, and the helpers of refinement
// Field points: DINCP(j, k) = grad(phi_I) . n at each hull panel (n x nbeta);
// Z = [W; -X] (2n x nrhs); the free-term test onb[nfp]; the finished values
// out[col * ld + p] (and inc[k * ld + p] when inc is not null).
// mirror >= 0: DINCP at the mirrored sources; flip: the modes whose generator
// changes sign (bit m for mode m + 1); accumulate: OR into onb instead of
// clearing it first.
void dincp(hipStream_t s, const MeshDev& m, const FreqDev& f, const WaveDev& w, double2* out, int mirror = -1);
void field_weights(hipStream_t s, const MeshDev& m, int nbody, int nbeta, const double2* dincp, const double2* X,
                   double2* Z, int flip = 0);
void on_body(hipStream_t s, const MeshDev& m, const double* vtx, int nfp, const double* px, const double* py,
             const double* pz, int* onb, bool accumulate = false);
// q = the mirror images of count points (mirror 0: x = 0, 1: y = 0)
void reflect_points(hipStream_t s, int count, int mirror, const double* px, const double* py, const double* pz,
                    double* qx, double* qy, double* qz);
void field_finish(hipStream_t s, int nfp, const double* px, const double* py, const double* pz, const int* onb,
                  const FreqDev& f, const WaveDev& w, int nd, int kind, const double2* P, size_t ld, double2* out,
                  double2* inc);
// Motion amplitudes of every heading (SolveMotionMulti): coef = the force
// reductions [amas | bdmp | exfc], bm = the five block-diagonal body
// matrices, nd = 6 nbody <= 192; work nbeta x nd x nd.
void motions(hipStream_t s, int nd, int nbeta, double w1, const double* coef, const double* bm, int quad,
             double2* work, double2* dspl, int* iters);
void to_c64(hipStream_t s, const double2* A, float2* Af, size_t count);
void to_c128(hipStream_t s, const float2* Af, double2* A, size_t count);
// C(r, c) = C(c, r) above the diagonal (n x n): completes a lower-triangle syrk
void mirror_lower(hipStream_t s, double2* C, int n);
void lu_factor_c64(hipStream_t s, float2* A, int n, int* ipiv, int* perm, int* work_idx, int* work_cnt, int* info,
                   double2* diag, const LuAux* aux = nullptr);
// count <= kLuBatchMax matrices of order n factored by one launch sequence
// (complex128, or complex64 with c64); each matrix's factors, pivots and INFO
// are those lu_factor / lu_factor_c64 give it alone, bit for bit.
void lu_factor_batch(hipStream_t s, const LuBatch& b, int count, int n, bool c64, const LuAux* aux = nullptr);
// skip (may be null, here and below): a device flag that turns every
// launch into a no-op when set (refinement after convergence)
void lu_solve_c64(hipStream_t s, const float2* LU, int n, const int* perm, double2* X, double2* Xtmp, double2* inv,
                  int nrhs, const int* skip = nullptr);
// Y = beta B + alpha A X (B may be null), A read once; P: matvec_workspace(n, nrhs) complex128
size_t matvec_workspace(int n, int nrhs);
void matvec(hipStream_t s, const double2* A, int n, const double2* X, int nrhs, const double2* B, double alpha,
            double beta, double2* Y, double2* P, const int* skip = nullptr);
// R = B - A X with per-column max norms of R (norms[0:nrhs]) and X (norms[nrhs:2 nrhs])
void residual(hipStream_t s, const double2* A, int n, const double2* X, const double2* B, double2* R, int nrhs,
              double* norms, double2* P, const int* skip = nullptr);
void axpy_update(hipStream_t s, double2* X, const double2* D, size_t count, const int* skip = nullptr);
// nsys independent LU-factored systems solved at once: system z of order
// nz[z] (nmax the largest) has factors at LU + lu_off[z], its permutation at
// perm[r0[z] .. r0[z+1]) (local indices) and owns rows [r0[z], r0[z+1]) of
// the n x nrhs block V (body[i] = system of row i); Z = the solutions in the
// same rows.  P, Pt: n x nrhs scratch; x_off[z] = nrhs * r0[z].
// inv / inv_off: the systems' inverted diagonal blocks, tri_inverse_len(nz[z])
// each, written by tri_inverse_systems once per factorisation.
void tri_inverse_systems(hipStream_t s, int nsys, int nmax, const double2* LU, const size_t* lu_off, const int* nz,
                         const size_t* inv_off, double2* inv);
void lu_solve_systems(hipStream_t s, int nsys, int nmax, const double2* LU, const size_t* lu_off, const int* nz,
                      const double2* inv, const size_t* inv_off, const size_t* x_off, const int* r0, const int* body,
                      const int* perm, int n, const double2* V, double2* Z, double2* P, double2* Pt, int nrhs);
// refinement state[0] = converged, state[1] = correction steps: reset, then
// after each residual test max_q |R_q|_max / |X_q|_max <= tol
void refine_reset(hipStream_t s, int* state);
// The refined complex64 solve (initial solve, up to iters residual /
// correction steps, final residual) of count <= kLuBatchMax systems of order
// n at once; each system's X, R, norms and state are what the same steps
// give it alone, bit for bit.
void refine_solve_batch(hipStream_t s, const RefineBatch& b, int count, int n, int nrhs, int iters, double tol);
void refine_check(hipStream_t s, const double* norms, int nrhs, double tol, int* state);
void gm_start(hipStream_t s, const GmDev& g, int nrhs, const double* norms);
void gm_restart(hipStream_t s, const GmDev& g, int nrhs, int m, const double* beta, double tol, double* scale,
                int* out, double* worst);
void gm_step(hipStream_t s, const GmDev& g, int nrhs, int m, int j, const double2* d1, const double2* d2,
             const double* hn, double tol, double* scale, int* out);
void gm_solve(hipStream_t s, const GmDev& g, int nrhs, int m, double2* y);
// restarted GMRES helpers (krylov.hip); blocks are n x nrhs, V holds nv of them
void block_dots(hipStream_t s, const double2* V, int nv, const double2* W, int n, int nrhs, double2* dots);
void block_update(hipStream_t s, const double2* V, int nv, double2* W, int n, int nrhs, const double2* h,
                  double2* hsum);
void col_norm2(hipStream_t s, const double2* W, int n, int nrhs, double* norms);
void col_scale(hipStream_t s, const double2* W, int n, int nrhs, const double* scale, double2* Y);
void block_combine(hipStream_t s, const double2* V, int m, int n, int nrhs, const double2* y, double2* U);
void rows_pack(hipStream_t s, const double2* X, int n, int r0, int nb, int nrhs, double2* P);
void rows_unpack(hipStream_t s, const double2* P, int n, int r0, int nb, int nrhs, double2* X);
void block_copy(hipStream_t s, const double2* A, int n, int r0, int nb, double2* B);
// H-matrix (hmatrix.hip): ACA of nblk admissible blocks of A (lda = n), the
// dense-block copies, and Y = beta B + alpha H X (Xp, Yp: n x nrhs scratch)
void hm_aca(hipStream_t s, const double2* A, int lda, const int* perm, int nblk, const int4* desc, const size_t* uoff,
            const size_t* voff, const size_t* roff, unsigned char* used, double2* U, double2* VH, int* rank,
            int rank_max, double tol);
void hm_dense_gather(hipStream_t s, const double2* A, int lda, const int* perm, int nblk, const int4* desc,
                     const size_t* doff, double2* D);
// c64: read the complex64 copies of the blocks (when built)
void hm_matvec(hipStream_t s, const HmDev& h, int n, const double2* X, int nrhs, const double2* B, double alpha,
               double beta, double2* Y, double2* Xp, double2* Yp, bool c64 = false);
}  // namespace launch

}  // namespace hamsx
