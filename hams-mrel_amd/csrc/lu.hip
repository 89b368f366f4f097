// lu.hip -- dense complex LU with partial (tournament) pivoting and the
// triangular solves, gfx950.  Replaces ZGETRF/ZGETRS (AssbMatxMulti.f90:112,
// :279, :346).
//
// Right-looking blocked factorisation, panel width NB = 32:
//   1. pivot selection for the m x NB panel by a tournament (CALU): 128-row
//      blocks are factored with partial pivoting in LDS, each keeps its NB
//      pivot rows, and groups of four candidate sets are merged the same way
//      until one set is left -- no grid-wide synchronisation;
//   2. the selected rows are swapped to the top across all N columns (the
//      swaps are also applied to a permutation vector used by the solves);
//   3. the NB x NB diagonal block is factored without further pivoting and
//      L21 = A21 U11^-1, U12 = L11^-1 A12 are formed row/column-parallel;
//   4. the trailing update A22 -= L21 U12 is a complex GEMM on MFMA
//      (v_mfma_f64_16x16x4f64 for complex128, v_mfma_f32_16x16x4f32 for
//      complex64), 64 x 64 output tiles per 256-thread block.
// The same templates serve the complex128 (plain) and complex64 (mixed
// precision) paths.
#include <hip/hip_runtime.h>

#include "hamsx_internal.hpp"

// wave-wide maximum from the ROCm device library (DPP reduction, result in
// every lane)
extern "C" __device__ __attribute__((const)) double __ockl_wfred_max_f64(double);

namespace hamsx {

template <typename R>
struct CT;
template <>
struct CT<double> {
  using c2 = double2;
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mfma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // accumulator element q of lane group g (= lane / 16) holds output row 4q + g
  static __device__ __forceinline__ int out_row(int g, int q) { return 4 * q + g; }
};
template <>
struct CT<float> {
  using c2 = float2;
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mfma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // accumulator element q of lane group g (= lane / 16) holds output row 4g + q
  static __device__ __forceinline__ int out_row(int g, int q) { return 4 * g + q; }
};

template <typename R>
__device__ __forceinline__ void crecip(R ar, R ai, R& rr, R& ri) {  // 1/(ar + i ai), Smith
  if (fabs(ar) >= fabs(ai)) {
    R q = ai / ar, d = ar + ai * q;
    rr = R(1) / d;
    ri = -q / d;
  } else {
    R q = ar / ai, d = ar * q + ai;
    rr = q / d;
    ri = R(-1) / d;
  }
}

// ---------------------------------------------------------------------------
// Tournament pivot selection.  One thread per candidate row, ROWS (256, 512
// or 1024) per block.  Level 0 (idx_in == nullptr): block b takes rows
// [row0 + b*ROWS, ...).  Level > 0: block b merges candidate sets
// [b*merge, b*merge + merge) of the previous level, merge = ROWS / NB.
// Larger blocks need fewer levels (at N = 3,968: three levels of 256-row
// blocks, two of 512) at a higher cost per step (more waves meet at each
// barrier); the choice is a tuning option (LuAux::sel_rows).  Each block runs
// partial pivoting on its candidates' panel entries (pivot = largest
// |re| + |im| as IZAMAX, ties to the first candidate).  Outputs the chosen
// rows in pivot order; the final level (FINAL, diag set) also writes the
// factored NB x NB diagonal block L\U in pivot order and ZGETRF's INFO.
//
// Each candidate's NB entries live in registers, slot k = column k, and the
// NB steps are unrolled so every slot index is static: step c touches only
// columns c.. (an earlier rolled form rotated the whole window by one slot
// per step to keep the pivot column in slot 0, which cost a register move
// per slot and an update of the retired slots every step).  One barrier per
// step: each wave finds its own largest |re| + |im| (a DPP wave reduction)
// and its first lane holding it (a ballot), and that lane publishes its
// value, row number and live columns to LDS; after the barrier every thread
// picks the winner among the waves and reads the pivot row from LDS.
// Buffers alternate between steps, so a step's writes never meet the
// previous step's reads.
// ---------------------------------------------------------------------------
constexpr int kOuter = 4 * kPanelNB;  // LU outer block: four panels share one far trailing update
// Outer blocks from this order on.  Measured (FP64, look-ahead): N = 3,968
// factors in 30.3 ms per panel and 34.0 ms in outer blocks (the far U chain
// at each block's end sits on the critical path of a selection-bound LU);
// N = 10,026 in 137 and 133 ms, and four solves in flight then run 4.4 and
// 5.1 solves/s (the far trailing matrix is read and written once per block).
constexpr int kOuterMinN = 6144;
static_assert((kPanelNB & (kPanelNB - 1)) == 0, "the rotating window indexes modulo NB");
// Turn the final pivot-row list (winners[c] = the row chosen at step c)
// into LAPACK-style sequential swaps ipiv[col0 + c] (1-based); one thread.
__device__ __forceinline__ void pivots_to_swaps(const int* winners, int* __restrict__ ipiv, int nsel, int col0) {
  int who[2 * kPanelNB], pos[2 * kPanelNB], nt = 0;  // rows whose position changed
  for (int c = 0; c < nsel; c++) {
    int w = winners[c], p = w;
    for (int q = 0; q < nt; q++)
      if (who[q] == w) {
        p = pos[q];
        break;
      }
    const int dst = col0 + c;
    ipiv[dst] = p + 1;
    if (p != dst) {
      // row currently at dst moves to p
      int rowat = dst;
      for (int q = 0; q < nt; q++)
        if (pos[q] == dst) {
          rowat = who[q];
          break;
        }
      bool f1 = false, f2 = false;
      for (int q = 0; q < nt; q++) {
        if (who[q] == w) {
          pos[q] = dst;
          f1 = true;
        } else if (who[q] == rowat) {
          pos[q] = p;
          f2 = true;
        }
      }
      if (!f1) {
        who[nt] = w;
        pos[nt++] = dst;
      }
      if (!f2) {
        who[nt] = rowat;
        pos[nt++] = p;
      }
    }
  }
}

// Batched: matrix blockIdx.y of bat; the tournament's candidate lists
// alternate between two buffers of wsets sets in the matrix's workspace
// (in_buf < 0: level 0, rows taken from the matrix itself).
template <typename R, bool FINAL, int ROWS>
__global__ __launch_bounds__(ROWS) void k_select(const LuBatch bat, int n, int col0, int ncols, int row0, int in_buf,
                                                int nsets_in, int out_buf, int wsets) {
  constexpr int kMerge = ROWS / kPanelNB, kSelWaves = ROWS / 64;
  const LuMat& mat = bat.m[blockIdx.y];
  const typename CT<R>::c2* __restrict__ A = static_cast<const typename CT<R>::c2*>(mat.A);
  const int lda = n;
  const int* __restrict__ idx_in = in_buf < 0 ? nullptr : mat.widx + (size_t)in_buf * wsets * kPanelNB;
  const int* __restrict__ cnt_in = in_buf < 0 ? nullptr : mat.wcnt + (size_t)in_buf * wsets;
  int* __restrict__ idx_out = mat.widx + (size_t)out_buf * wsets * kPanelNB;
  int* __restrict__ cnt_out = mat.wcnt + (size_t)out_buf * wsets;
  typename CT<R>::c2* __restrict__ diag = static_cast<typename CT<R>::c2*>(mat.diag);
  int* __restrict__ info = mat.info;
  __shared__ int sidx[ROWS];
  __shared__ int off[kMerge + 1];
  __shared__ double wmax[2][kSelWaves];
  __shared__ int wrow[2][kSelWaves];
  __shared__ double2 wwin[2][kSelWaves][kPanelNB];  // each wave's candidate pivot window
  __shared__ int sel[kPanelNB];
  const int t = threadIdx.x, b = blockIdx.x;
  int mc;
  if (idx_in == nullptr) {
    const int r0 = row0 + b * ROWS;
    mc = min(ROWS, n - r0);
    if (t < mc) sidx[t] = r0 + t;
  } else {
    if (t == 0) {
      int o = 0;
      for (int q = 0; q < kMerge; q++) {
        off[q] = o;
        const int si = b * kMerge + q;
        o += (si < nsets_in) ? cnt_in[si] : 0;
      }
      off[kMerge] = o;
    }
    __syncthreads();
    mc = off[kMerge];
    for (int q = 0; q < kMerge; q++) {
      const int si = b * kMerge + q, c = off[q + 1] - off[q];
      if (t < c) sidx[off[q] + t] = idx_in[si * kPanelNB + t];
    }
  }
  __syncthreads();
  bool active = t < mc;
  double vr[kPanelNB], vi[kPanelNB];  // computed in FP64 for either factor type
  {
    // unconditional loads at clamped, in-bounds addresses (all issued before
    // the first use), masked afterwards; columns >= ncols are zero
    const size_t g = (size_t)sidx[active ? t : 0];
    typename CT<R>::c2 v[kPanelNB];
#pragma unroll
    for (int q = 0; q < kPanelNB; q++) v[q] = A[g + (size_t)lda * (col0 + min(q, ncols - 1))];
#pragma unroll
    for (int q = 0; q < kPanelNB; q++) {
      const bool use = active && q < ncols;
      vr[q] = use ? (double)v[q].x : 0.0;
      vi[q] = use ? (double)v[q].y : 0.0;
    }
  }
  const int nsel = __builtin_amdgcn_readfirstlane(min(ncols, mc));
  int mystep = -1, zero_at = 0;
  const int wave = t >> 6, lane = t & 63;
  // The steps are unrolled, so slot k holds column k throughout: step c
  // publishes, reads and updates only the live columns c+1.. (the pivot
  // column c itself is read for the multiplier), and the multiplier
  // replaces column c in place.
#pragma unroll
  for (int c = 0; c < kPanelNB; c++) {
    if (c >= nsel) continue;  // block-uniform; a constant trip count lets the loop unroll
    const int buf = c & 1;
    const double m = active ? fabs(vr[c]) + fabs(vi[c]) : -1.0;
    const double mw = __ockl_wfred_max_f64(m);
    // ties go to the lowest row; with NaN in every lane no lane matches and
    // lane 0 answers, so each wave always publishes a valid row
    const unsigned long long hit = __ballot(m == mw);
    const int first = hit ? __builtin_ctzll(hit) : 0;
    if (lane == first) {
#pragma unroll
      for (int k = c; k < kPanelNB; k++) wwin[buf][wave][k] = make_double2(vr[k], vi[k]);
      wmax[buf][wave] = mw;
      wrow[buf][wave] = t;
    }
    __syncthreads();
    double best = wmax[buf][0];
    int w = 0;
    for (int q = 1; q < kSelWaves; q++)
      if (wmax[buf][q] > best) {  // equal maxima keep the lower wave, i.e. the lower row
        best = wmax[buf][q];
        w = q;
      }
    const int p = wrow[buf][w];
    const bool nonzero = best > 0.0;
    if (t == p) {
      sel[c] = t;
      mystep = c;
      active = false;
    }
    if (active && nonzero) {
      const double2 pv = wwin[buf][w][c];
      double ir = 0.0, ii = 0.0;
      crecip(pv.x, pv.y, ir, ii);
      // the fused multiply-adds are written out, so the rounding does not
      // depend on how the compiler contracts the expression in each step
      const double lr = fma(vr[c], ir, -(vi[c] * ii));
      const double li = fma(vi[c], ir, vr[c] * ii);
#pragma unroll
      for (int k = c + 1; k < kPanelNB; k++) {
        const double2 u = wwin[buf][w][k];
        vr[k] -= fma(lr, u.x, -(li * u.y));
        vi[k] -= fma(li, u.x, lr * u.y);
      }
      vr[c] = lr;
      vi[c] = li;
    }
    if (!nonzero && zero_at == 0) zero_at = c + 1;
  }
  __syncthreads();
  if (t < nsel) idx_out[b * kPanelNB + t] = sidx[sel[t]];
  if (t == 0) cnt_out[b] = nsel;
  if (FINAL) {
    // the swaps of this panel (the final level is one block per matrix)
    __shared__ int win[kPanelNB];
    if (t < nsel) win[t] = sidx[sel[t]];
    __syncthreads();
    if (t == 0) pivots_to_swaps(win, mat.ipiv, nsel, col0);
    // the pivot row of step s: the multipliers of steps 0..s-1 in slots
    // 0..s-1 (row s of L) and row s of U in slots s..
    if (mystep >= 0)
#pragma unroll
      for (int k = 0; k < kPanelNB; k++)
        if (k < ncols) diag[mystep + kPanelNB * k] = {(R)vr[k], (R)vi[k]};
    if (t == 0 && zero_at && info[0] == 0) info[0] = col0 + zero_at;
  }
}

// Apply the swaps of pivot rows [col0, col0 + nsel) in order to the columns
// [a0, a1) and [b0, b1) (one thread per column) and, with do_perm, to the
// permutation vector (one more thread).
template <typename R>
__global__ void k_swap_rows(const LuBatch bat, int n, int col0, int nsel, int a0, int a1, int b0, int b1,
                            int do_perm) {
  const LuMat& mat = bat.m[blockIdx.y];
  typename CT<R>::c2* __restrict__ A = static_cast<typename CT<R>::c2*>(mat.A);
  const int lda = n;
  const int* __restrict__ ipiv = mat.ipiv;
  int* __restrict__ perm = mat.perm;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x, na = a1 - a0, nb_ = b1 - b0;
  int j;
  if (tid < na) j = a0 + tid;
  else if (tid < na + nb_) j = b0 + (tid - na);
  else if (do_perm && tid == na + nb_) j = n;
  else return;
  for (int c = 0; c < nsel; c++) {
    const int r = col0 + c, p = ipiv[r] - 1;
    if (p == r) continue;
    if (j < n) {
      auto a = A[(size_t)r + (size_t)lda * j];
      A[(size_t)r + (size_t)lda * j] = A[(size_t)p + (size_t)lda * j];
      A[(size_t)p + (size_t)lda * j] = a;
    } else {
      int q = perm[r];
      perm[r] = perm[p];
      perm[p] = q;
    }
  }
}

// With the diagonal block factored: L21 = A21 U11^-1 (blocks [0, nrb), one
// row per thread) and U12 = L11^-1 A12 for the columns [uc0, uc1) (blocks
// [nrb, ..), one column per thread).  The factored diagonal block is read
// from the matrix's diagonal scratch (from the final tournament level), or
// with diag_in_A from A itself at (col0, col0); with store set, block 0 also
// stores it into A.  No block reads what another writes.
template <typename R>
__global__ __launch_bounds__(256) void k_panel_solve(const LuBatch bat, int n, int col0, int nb, int nrb,
                                                     int diag_in_A, int store, int uc0, int uc1) {
  const LuMat& mat = bat.m[blockIdx.y];
  typename CT<R>::c2* __restrict__ A = static_cast<typename CT<R>::c2*>(mat.A);
  const int lda = n;
  const typename CT<R>::c2* __restrict__ diag =
      diag_in_A ? A + (size_t)col0 + (size_t)n * col0 : static_cast<const typename CT<R>::c2*>(mat.diag);
  const int ldd = diag_in_A ? n : kPanelNB;
  __shared__ R dr[kPanelNB][kPanelNB + 1];
  __shared__ R di[kPanelNB][kPanelNB + 1];
  __shared__ R rr[kPanelNB], ri[kPanelNB];  // reciprocals of the pivots
  const int t = threadIdx.x;
  for (int e = t; e < nb * nb; e += 256) {
    int r = e % nb, c = e / nb;
    auto v = diag[r + (size_t)ldd * c];
    dr[r][c] = v.x;
    di[r][c] = v.y;
    if (store && blockIdx.x == 0) A[(size_t)(col0 + r) + (size_t)lda * (col0 + c)] = v;
  }
  __syncthreads();
  if (t < nb) {
    R a = R(0), b = R(0);
    if (dr[t][t] != R(0) || di[t][t] != R(0)) crecip(dr[t][t], di[t][t], a, b);
    rr[t] = a;
    ri[t] = b;
  }
  __syncthreads();
  const int b = blockIdx.x;
  if (b < nrb) {
    const int i = col0 + nb + b * 256 + t;
    if (i >= n) return;
    R xr[kPanelNB], xi[kPanelNB];
#pragma unroll
    for (int c = 0; c < kPanelNB; c++) {  // unconditional, clamped: all loads in flight together
      auto v = A[(size_t)i + (size_t)lda * (col0 + min(c, nb - 1))];
      xr[c] = v.x;
      xi[c] = v.y;
    }
#pragma unroll
    for (int c = 0; c < kPanelNB; c++) {
      if (c < nb) {
        R ar = xr[c], ai = xi[c];
#pragma unroll
        for (int q = 0; q < kPanelNB; q++) {
          if (q < c) {
            ar -= xr[q] * dr[q][c] - xi[q] * di[q][c];
            ai -= xr[q] * di[q][c] + xi[q] * dr[q][c];
          }
        }
        xr[c] = ar * rr[c] - ai * ri[c];
        xi[c] = ar * ri[c] + ai * rr[c];
      }
    }
#pragma unroll
    for (int c = 0; c < kPanelNB; c++)
      if (c < nb) A[(size_t)i + (size_t)lda * (col0 + c)] = {xr[c], xi[c]};
  } else {
    const int j = uc0 + (b - nrb) * 256 + t;
    if (j >= uc1) return;
    R yr[kPanelNB], yi[kPanelNB];
#pragma unroll
    for (int c = 0; c < kPanelNB; c++) {
      auto v = A[(size_t)(col0 + min(c, nb - 1)) + (size_t)lda * j];
      yr[c] = v.x;
      yi[c] = v.y;
    }
#pragma unroll
    for (int c = 0; c < kPanelNB; c++) {
      if (c < nb) {
#pragma unroll
        for (int q = 0; q < kPanelNB; q++) {
          if (q < c) {
            yr[c] -= dr[c][q] * yr[q] - di[c][q] * yi[q];
            yi[c] -= dr[c][q] * yi[q] + di[c][q] * yr[q];
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < kPanelNB; c++)
      if (c < nb) A[(size_t)(col0 + c) + (size_t)lda * j] = {yr[c], yi[c]};
  }
}

// Update C -= L U of the rows [ib, ie) and the column tiles tn0, tn0 + 1, ...
// (64 columns each, from column jb, ending at je), L = A[:, kb .. kb + kk),
// U = A[kb .. kb + kk, :]; ntiles = tiles_m x the number of column tiles.  MFMA 16x16x4 computes the transposed product
// (U^T L^T) so that the 16 consecutive lanes of an accumulator row map to
// 16 consecutive rows of column-major C.  Operand layout (both types): lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16]; the accumulator row held
// by element q of lane l is type-specific (CT::out_row, measured on gfx950).  Block tiles are 64 x 64, each of
// the four waves owns a 32 x 32 quarter (2 x 2 MFMA tiles, real and
// imaginary accumulators).  Block ids are remapped so that each of the eight
// XCDs works on a contiguous band of tiles (L21 rows stay in its L2).
template <typename R>
__global__ __launch_bounds__(256) void k_trailing_update(const LuBatch bat, int lda, int ib, int ie, int jb, int je,
                                                         int kb, int kk, int tiles_m, int ntiles, int tn0) {
  using v4 = typename CT<R>::v4;
  typename CT<R>::c2* __restrict__ A = static_cast<typename CT<R>::c2*>(bat.m[blockIdx.y].A);
  __shared__ R lr_[kPanelNB][64], li_[kPanelNB][64];  // L tile: [k][row]
  __shared__ R ur_[kPanelNB][64], ui_[kPanelNB][64];  // U tile: [k][col]
  int bid = blockIdx.x;
  {
    const int per = (ntiles + 7) / 8, xcd = bid % 8, k = bid / 8;
    const int rb = xcd * per + k;
    if (rb >= ntiles) return;
    bid = rb;
  }
  const int tm = bid % tiles_m, tn = tn0 + bid / tiles_m;
  const int i0 = ib + tm * 64, j0 = jb + tn * 64;
  const int t = threadIdx.x;
  const int w = t >> 6, lane = t & 63;
  const int wi = (w & 1) * 32, wj = (w >> 1) * 32;  // wave quarter within the tile
  const int l16 = lane & 15, kq = lane >> 4;
  v4 accR[2][2], accI[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      accR[a][b] = v4{0, 0, 0, 0};
      accI[a][b] = v4{0, 0, 0, 0};
    }
  // K in chunks of kPanelNB through the same LDS tiles; the accumulators
  // carry the whole sum, so C is read and written once
  for (int kc0 = 0; kc0 < kk; kc0 += kPanelNB) {
  const int kn = min(kPanelNB, kk - kc0);
  if (kc0 > 0) __syncthreads();  // every wave is done with the previous chunk
  {
    // unconditional loads at clamped, in-bounds addresses, all in flight
    // together; out-of-range entries are zeroed when staged to LDS
    constexpr int kPer = 64 * kPanelNB / 256;
    typename CT<R>::c2 lv[kPer], uv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      const int e = t + 256 * u, r = e % 64, k = e / 64;
      const int kc = kb + kc0 + min(k, kn - 1);
      lv[u] = A[(size_t)min(i0 + r, ie - 1) + (size_t)lda * kc];
      uv[u] = A[(size_t)kc + (size_t)lda * min(j0 + r, je - 1)];
    }
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      const int e = t + 256 * u, r = e % 64, k = e / 64;
      const bool lk = k < kn && i0 + r < ie, uk = k < kn && j0 + r < je;
      lr_[k][r] = lk ? lv[u].x : R(0);
      li_[k][r] = lk ? lv[u].y : R(0);
      ur_[k][r] = uk ? uv[u].x : R(0);
      ui_[k][r] = uk ? uv[u].y : R(0);
    }
  }
  __syncthreads();
  for (int k0 = 0; k0 < kn; k0 += 4) {
    const int k = k0 + kq;
    R Ur[2], Ui[2], Lr[2], Li[2];
    for (int b = 0; b < 2; b++) {  // A operand: U^T rows = C columns
      Ur[b] = ur_[k][wj + b * 16 + l16];
      Ui[b] = ui_[k][wj + b * 16 + l16];
    }
    for (int a = 0; a < 2; a++) {  // B operand: L^T columns = C rows
      Lr[a] = lr_[k][wi + a * 16 + l16];
      Li[a] = li_[k][wi + a * 16 + l16];
    }
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        accR[a][b] = CT<R>::mfma(Ur[b], Lr[a], accR[a][b]);
        accR[a][b] = CT<R>::mfma(-Ui[b], Li[a], accR[a][b]);
        accI[a][b] = CT<R>::mfma(Ui[b], Lr[a], accI[a][b]);
        accI[a][b] = CT<R>::mfma(Ur[b], Li[a], accI[a][b]);
      }
  }
  }
  // read the 16 C entries first (clamped addresses, all loads in flight),
  // then write back the in-range ones
  typename CT<R>::c2 cv[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = min(i0 + wi + a * 16 + l16, ie - 1);
        const int col = min(j0 + wj + b * 16 + CT<R>::out_row(kq, q), je - 1);
        cv[a][b][q] = A[(size_t)row + (size_t)lda * col];
      }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = i0 + wi + a * 16 + l16;
        const int col = j0 + wj + b * 16 + CT<R>::out_row(kq, q);
        if (row < ie && col < je) {
          auto v = cv[a][b][q];
          v.x -= accR[a][b][q];
          v.y -= accI[a][b][q];
          A[(size_t)row + (size_t)lda * col] = v;
        }
      }
}

__global__ void k_lu_init(const LuBatch bat, int n) {
  int* __restrict__ perm = bat.m[blockIdx.y].perm;
  int* __restrict__ info = bat.m[blockIdx.y].info;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) perm[i] = i;
  if (i == 0) info[0] = 0;
}

// One tournament level: `blocks` selection blocks of `rows` threads per matrix.
template <typename R>
static void select_level(hipStream_t s, int rows, bool final, int blocks, const LuBatch& bat, int nbat, int n,
                         int col0, int nb, int in_buf, int nsets_in, int out_buf, int wsets) {
#define HX_SEL(ROWS)                                                                                              \
  if (final)                                                                                                     \
    HX_LAUNCH((k_select<R, true, ROWS>), dim3(blocks, nbat), dim3(ROWS), 0, s, bat, n, col0, nb, col0, in_buf,    \
              nsets_in, out_buf, wsets);                                                                         \
  else                                                                                                           \
    HX_LAUNCH((k_select<R, false, ROWS>), dim3(blocks, nbat), dim3(ROWS), 0, s, bat, n, col0, nb, col0, in_buf,   \
              nsets_in, out_buf, wsets);
  if (rows == 1024) {
    HX_SEL(1024)
  } else if (rows == 512) {
    HX_SEL(512)
  } else {
    HX_SEL(kSelRows)
  }
#undef HX_SEL
}

// Factors the nbat matrices of bat (all of order n) with one launch sequence:
// every launch covers all of them (grid.y = matrix), so each step's latency
// chain serves the whole batch.  Each matrix gets exactly the operations of a
// factorisation on its own.
template <typename R>
static void lu_factor_t(hipStream_t s, const LuBatch& bat, int nbat, int n, const LuAux* aux) {
  HX_LAUNCH(k_lu_init, dim3((n + 255) / 256, nbat), dim3(256), 0, s, bat, n);
  const bool la = aux && aux->s2;
  const int rows = aux ? aux->sel_rows : kSelRows;
  const int outer_min = aux ? aux->outer_min_n : kOuterMinN;
  bool pending = false;  // a deferred trailing update is running on aux->s2
  // workspace: two ping-pong candidate buffers of nsets0 sets (sized for
  // kSelRows-row blocks, the smallest)
  const int nsets0 = (n + rows - 1) / rows, merge = rows / kPanelNB;
  for (int col0 = 0; col0 < n; col0 += kPanelNB) {
    const int nb = n - col0 < kPanelNB ? n - col0 : kPanelNB;
    const int m = n - col0;
    int nsets = (m + rows - 1) / rows, cur = 0;
    select_level<R>(s, rows, nsets == 1, nsets, bat, nbat, n, col0, nb, -1, 0, cur, nsets0);
    while (nsets > 1) {
      const int nout = (nsets + merge - 1) / merge;
      select_level<R>(s, rows, nout == 1, nout, bat, nbat, n, col0, nb, cur, nsets, 1 - cur, nsets0);
      cur = 1 - cur;
      nsets = nout;
    }
    // Outer blocks of kOuter columns: within the block each panel updates only
    // the block's columns; at the block's end the columns to its right get
    // U = L11^-1 A12 for all of the block's rows and one rank-kOuter update,
    // so the far trailing matrix is read and written once per block instead
    // of once per panel.
    const int rest = n - col0 - nb;
    const bool outer = !(outer_min <= 0 || n < outer_min);
    if (!outer) {
      // the row swaps touch every column: the previous step's deferred update must be done
      if (pending) (void)hipStreamWaitEvent(s, aux->update, 0);
      pending = false;
      HX_LAUNCH(k_swap_rows<R>, dim3((n + 1 + 255) / 256, nbat), dim3(256), 0, s, bat, n, col0, nb, 0, n, 0, 0, 1);
    }
    if (!outer) {  // one panel at a time: U12 and the update over every trailing column
      const int nrb = (rest + 255) / 256;
      HX_LAUNCH(k_panel_solve<R>, dim3(nrb > 0 ? nrb * 2 : 1, nbat), dim3(256), 0, s, bat, n, col0, nb, nrb, 0, 1,
                col0 + nb, n);
      if (rest > 0) {
        const int tiles = (rest + 63) / 64;
        if (!la || tiles == 1) {
          const int ntiles = tiles * tiles;
          HX_LAUNCH(k_trailing_update<R>, dim3(((ntiles + 7) / 8) * 8, nbat), dim3(256), 0, s, bat, n, col0 + nb,
                    n, col0 + nb, n, col0, nb, tiles, ntiles, 0);
        } else {
          // look-ahead: the first 64 trailing columns (the next panel) here, the
          // rest on the second stream, overlapping the next pivot selection,
          // which reads only the next panel's columns
          HX_LAUNCH(k_trailing_update<R>, dim3(((tiles + 7) / 8) * 8, nbat), dim3(256), 0, s, bat, n, col0 + nb,
                    n, col0 + nb, n, col0, nb, tiles, tiles, 0);
          (void)hipEventRecord(aux->panel, s);
          (void)hipStreamWaitEvent(aux->s2, aux->panel, 0);
          const int ntiles = tiles * (tiles - 1);
          HX_LAUNCH(k_trailing_update<R>, dim3(((ntiles + 7) / 8) * 8, nbat), dim3(256), 0, aux->s2, bat, n,
                    col0 + nb, n, col0 + nb, n, col0, nb, tiles, ntiles, 1);
          (void)hipEventRecord(aux->update, aux->s2);
          pending = true;
        }
      }
      continue;
    }
    const int blk0 = col0 - col0 % kOuter, blk1 = min(blk0 + kOuter, n);
    // The panel's swaps on the block's own columns (and the permutation)
    // only: the columns left and right of the block take the block's swaps
    // at its end, so the next block's panels never wait for the far update
    // of this one, which runs on the second stream meanwhile (and reads the
    // block's L columns, which these swaps do not touch).
    HX_LAUNCH(k_swap_rows<R>, dim3((blk1 - blk0 + 1 + 255) / 256, nbat), dim3(256), 0, s, bat, n, col0, nb, blk0,
              blk1, 0, 0, 1);
    const int near = blk1 - col0 - nb;
    const int nrb = (rest + 255) / 256, ncb = (near + 255) / 256;
    HX_LAUNCH(k_panel_solve<R>, dim3(nrb + ncb > 0 ? nrb + ncb : 1, nbat), dim3(256), 0, s, bat, n, col0, nb, nrb, 0,
              1, col0 + nb, blk1);
    if (near > 0) {
      const int tm = (rest + 63) / 64, tn = (near + 63) / 64, ntiles = tm * tn;
      HX_LAUNCH(k_trailing_update<R>, dim3(((ntiles + 7) / 8) * 8, nbat), dim3(256), 0, s, bat, n, col0 + nb, n,
                col0 + nb, blk1, col0, nb, tm, ntiles, 0);
    }
    if (col0 + nb == blk1) {
      // the block's swaps on every other column, once the previous block's
      // far update (which covers this block's far columns) is done
      if (pending) (void)hipStreamWaitEvent(s, aux->update, 0);
      pending = false;
      const int other = n - (blk1 - blk0);
      if (other > 0)
        HX_LAUNCH(k_swap_rows<R>, dim3((other + 255) / 256, nbat), dim3(256), 0, s, bat, n, blk0, blk1 - blk0, 0,
                  blk0, blk1, n, 0);
    }
    if (col0 + nb == blk1 && blk1 < n) {
      const int far = n - blk1, tm = (far + 63) / 64;
      // U = L11^-1 A12 of the block's rows over the far columns [c0, c1),
      // one panel's rows after another
      auto u_far = [&](hipStream_t st, int c0, int c1) {
        const int tn = (c1 - c0 + 63) / 64, fcb = (c1 - c0 + 255) / 256;
        for (int q0 = blk0; q0 < blk1; q0 += kPanelNB) {
          const int qnb = min(kPanelNB, blk1 - q0);
          if (q0 > blk0)
            HX_LAUNCH(k_trailing_update<R>, dim3(((tn + 7) / 8) * 8, nbat), dim3(256), 0, st, bat, n, q0, q0 + qnb,
                      c0, c1, blk0, q0 - blk0, 1, tn, 0);
          HX_LAUNCH(k_panel_solve<R>, dim3(fcb, nbat), dim3(256), 0, st, bat, n, q0, qnb, 0, 1, 0, c0, c1);
        }
      };
      // (U of all far columns stays on this stream: the second stream carries
      // the far update, which bounds the early blocks, and a U chain there
      // measured slower)
      u_far(s, blk1, n);
      if (!la || tm <= 2) {
        HX_LAUNCH(k_trailing_update<R>, dim3(((tm * tm + 7) / 8) * 8, nbat), dim3(256), 0, s, bat, n, blk1, n, blk1,
                  n, blk0, blk1 - blk0, tm, tm * tm, 0);
      } else {
        // look-ahead: the rank-kOuter update of the next outer block's
        // columns (two tiles) here, the rest on the second stream,
        // overlapping the next block's panels, which touch only their own
        // columns
        HX_LAUNCH(k_trailing_update<R>, dim3(((2 * tm + 7) / 8) * 8, nbat), dim3(256), 0, s, bat, n, blk1, n, blk1,
                  n, blk0, blk1 - blk0, tm, 2 * tm, 0);
        (void)hipEventRecord(aux->panel, s);
        (void)hipStreamWaitEvent(aux->s2, aux->panel, 0);
        const int ntiles = tm * (tm - 2);
        HX_LAUNCH(k_trailing_update<R>, dim3(((ntiles + 7) / 8) * 8, nbat), dim3(256), 0, aux->s2, bat, n, blk1, n,
                  blk1, n, blk0, blk1 - blk0, tm, ntiles, 2);
        (void)hipEventRecord(aux->update, aux->s2);
        pending = true;
      }
    }
  }
  if (pending) (void)hipStreamWaitEvent(s, aux->update, 0);
}

// ---------------------------------------------------------------------------
// Triangular solves.  X holds the right-hand sides (complex128); the factors
// are complex128 or complex64 (mixed path: FP32 factors, FP64 right-hand
// sides); arithmetic is FP64.  kTrsmRhs right-hand sides per workgroup.
// ---------------------------------------------------------------------------
constexpr int kTrsmRhs = 8;
// ---------------------------------------------------------------------------
// One factorisation, 256 rows per launch.  The 32 x 32 diagonal blocks of L
// (unit) and U are inverted once per solve (k_tri_inverse: substitution on
// the identity, a half-wave per column, lane r holding row r); a
// launch then takes a kSuper-row block of rows: its right-hand sides sit in
// LDS, each 32-row diagonal block is solved as a small product with its
// inverse and eliminated from the block's other rows, and every workgroup
// finally updates its 256 rows outside the block with the block's solution.
// Each workgroup solves the block for itself (nothing but the launch order
// synchronises workgroups), so the chain is n / 256 launches per triangle
// instead of n / 32.
// ---------------------------------------------------------------------------
constexpr int kSuper = 256;
static_assert(kSuper / kPanelNB * kPanelNB == kSuper && 256 / kPanelNB == kTrsmRhs,
              "thread (r, q) of a 256-thread block covers one row and one right-hand side of a diagonal block");

// inv[(2 b + u) * 1024 + c * 32 + r] = (T_bb^-1)(r, c): u = 0 the unit lower
// factor, u = 1 the upper factor; rows and columns past n are zero.
template <typename R>
__device__ __forceinline__ void tri_inverse_body(const typename CT<R>::c2* __restrict__ LU, int n,
                                                 double2* __restrict__ inv, int b, int upper) {
  __shared__ double tr[kPanelNB][kPanelNB], ti[kPanelNB][kPanelNB];  // [c][r]
  __shared__ double dr[kPanelNB], di[kPanelNB];
  const int t = threadIdx.x;
  const int k0 = b * kPanelNB, kb = min(kPanelNB, n - k0);
  for (int e = t; e < kPanelNB * kPanelNB; e += 256) {
    const int r = e % kPanelNB, c = e / kPanelNB;
    double vr = 0.0, vi = 0.0;
    if (r < kb && c < kb) {
      auto v = LU[(size_t)(k0 + r) + (size_t)n * (k0 + c)];
      vr = (double)v.x;
      vi = (double)v.y;
    }
    tr[c][r] = vr;
    ti[c][r] = vi;
  }
  __syncthreads();
  if (upper && t < kb) {  // reciprocal of the diagonal (Smith)
    double a = tr[t][t], bb = ti[t][t];
    if (fabs(a) >= fabs(bb)) {
      double qq = bb / a, dd = a + bb * qq;
      dr[t] = 1.0 / dd;
      di[t] = -qq / dd;
    } else {
      double qq = a / bb, dd = a * qq + bb;
      dr[t] = qq / dd;
      di[t] = -1.0 / dd;
    }
  }
  __syncthreads();
  const int lane = t & 63, half = lane >> 5, r = lane & 31;
  double2* out = inv + (size_t)(2 * b + upper) * kPanelNB * kPanelNB;
  for (int pass = 0; pass < kPanelNB / 8; pass++) {
    const int col = pass * 8 + (t >> 6) * 2 + half;  // the identity column this half-wave solves for
    double vr = (r == col && r < kb) ? 1.0 : 0.0, vi = 0.0;
    if (!upper) {
      for (int c = 0; c < kb; c++) {
        const double cr = __shfl(vr, (half << 5) + c), ci = __shfl(vi, (half << 5) + c);
        if (r > c && r < kb) {
          const double a = tr[c][r], bb = ti[c][r];
          vr -= a * cr - bb * ci;
          vi -= a * ci + bb * cr;
        }
      }
    } else {
      for (int c = kb - 1; c >= 0; c--) {
        if (r == c) {
          const double a = vr, bb = vi;
          vr = a * dr[c] - bb * di[c];
          vi = a * di[c] + bb * dr[c];
        }
        const double cr = __shfl(vr, (half << 5) + c), ci = __shfl(vi, (half << 5) + c);
        if (r < c) {
          const double a = tr[c][r], bb = ti[c][r];
          vr -= a * cr - bb * ci;
          vi -= a * ci + bb * cr;
        }
      }
    }
    out[(size_t)col * kPanelNB + r] = (r < kb && col < kb) ? make_double2(vr, vi) : make_double2(0.0, 0.0);
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_tri_inverse(const typename CT<R>::c2* __restrict__ LU, int n,
                                                     double2* __restrict__ inv, const int* __restrict__ skip) {
  if (skip && *skip) return;
  tri_inverse_body<R>(LU, n, inv, blockIdx.x, blockIdx.y);
}

// The same for a batch of systems (blockIdx.z = system z of order nz[z],
// factors at LU + lu_off[z], inverses at inv + inv_off[z]).
template <typename R>
__global__ __launch_bounds__(256) void k_tri_inverse_batched(const typename CT<R>::c2* __restrict__ LU,
                                                             const size_t* __restrict__ lu_off,
                                                             const int* __restrict__ nz,
                                                             const size_t* __restrict__ inv_off,
                                                             double2* __restrict__ inv) {
  const int z = blockIdx.z, n = nz[z];
  if ((int)blockIdx.x * kPanelNB >= n) return;
  tri_inverse_body<R>(LU + lu_off[z], n, inv + inv_off[z], blockIdx.x, blockIdx.y);
}

// One kSuper-row block [k0, k0 + kb) for the right-hand sides of group
// blockIdx.y: solves it (LOWER: forward, unit L; else backward, U) and
// updates this workgroup's 256 rows outside it (below it, or above it).
// The block's solution goes to Y (workgroup 0 of the group); the updated
// rows stay in X, which no workgroup reads inside the block.
template <typename R, bool LOWER>
__device__ __forceinline__ void trsm_super_body(const typename CT<R>::c2* __restrict__ LU,
                                                const double2* __restrict__ inv, int n, int k0, int kb,
                                                double2* __restrict__ X, double2* __restrict__ Y, int nrhs, int bx) {
  // [q][row]: a thread's own row is contiguous across lanes, a shared row is
  // one broadcast address
  __shared__ double2 xs[kTrsmRhs][kSuper];        // the block's right-hand sides, then its solution
  __shared__ double2 iv[kPanelNB][kPanelNB + 1];  // inverse of the current diagonal block, [c][r]
  const int t = threadIdx.x, q0 = blockIdx.y * kTrsmRhs;
  const int nq = min(kTrsmRhs, nrhs - q0);
  for (int e = t; e < kSuper * kTrsmRhs; e += 256) {
    const int r = e % kSuper, q = e / kSuper;
    xs[q][r] = (r < kb && q < nq) ? X[(size_t)(k0 + r) + (size_t)n * (q0 + q)] : make_double2(0.0, 0.0);
  }
  const int nsub = (kb + kPanelNB - 1) / kPanelNB;
  const int rr = t % kPanelNB, qq = t / kPanelNB;  // (row, right-hand side) of a diagonal block
  for (int ss = 0; ss < nsub; ss++) {
    const int sblk = LOWER ? ss : nsub - 1 - ss;
    const int r0 = sblk * kPanelNB, sb = min(kPanelNB, kb - r0);
    const double2* src = inv + (size_t)(2 * ((k0 + r0) / kPanelNB) + (LOWER ? 0 : 1)) * kPanelNB * kPanelNB;
    for (int e = t; e < kPanelNB * kPanelNB; e += 256) iv[e / kPanelNB][e % kPanelNB] = src[e];
    // this thread's factor row against the diagonal block (rows still to
    // come), all loads issued before the arithmetic
    const bool upd = LOWER ? (t >= r0 + sb && t < kb) : (t < r0);
    typename CT<R>::c2 l[kPanelNB];
#pragma unroll
    for (int c = 0; c < kPanelNB; c++)
      l[c] = LU[(size_t)(k0 + (upd ? t : r0)) + (size_t)n * (k0 + r0 + min(c, sb - 1))];
    __syncthreads();
    double yr = 0.0, yi = 0.0;
    for (int c = 0; c < sb; c++) {
      const double2 a = iv[c][rr], x = xs[qq][r0 + c];
      yr += a.x * x.x - a.y * x.y;
      yi += a.x * x.y + a.y * x.x;
    }
    __syncthreads();  // every read of the block's old values is done
    if (rr < sb) {
      xs[qq][r0 + rr] = make_double2(yr, yi);
      if (bx == 0 && qq < nq) Y[(size_t)(k0 + r0 + rr) + (size_t)n * (q0 + qq)] = make_double2(yr, yi);
    }
    __syncthreads();
    // eliminate the solved rows from the block's rows still to come
    if (upd) {
      double ar[kTrsmRhs], ai[kTrsmRhs];
#pragma unroll
      for (int q = 0; q < kTrsmRhs; q++) ar[q] = ai[q] = 0.0;
#pragma unroll
      for (int c = 0; c < kPanelNB; c++) {
        if (c < sb) {
          const double lr = (double)l[c].x, li = (double)l[c].y;
#pragma unroll
          for (int q = 0; q < kTrsmRhs; q++) {
            const double2 y = xs[q][r0 + c];
            ar[q] += lr * y.x - li * y.y;
            ai[q] += lr * y.y + li * y.x;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < kTrsmRhs; q++) xs[q][t] = make_double2(xs[q][t].x - ar[q], xs[q][t].y - ai[q]);
    }
    __syncthreads();
  }
  // this workgroup's rows outside the block, the factor row in chunks of 32
  // columns with each chunk's loads issued before its arithmetic
  const int i = LOWER ? k0 + kb + bx * 256 + t : bx * 256 + t;
  if (LOWER ? i >= n : i >= k0) return;
  double ar[kTrsmRhs], ai[kTrsmRhs];
#pragma unroll
  for (int q = 0; q < kTrsmRhs; q++) ar[q] = ai[q] = 0.0;
  for (int c0 = 0; c0 < kb; c0 += kPanelNB) {
    const int cb = min(kPanelNB, kb - c0);
    typename CT<R>::c2 l[kPanelNB];
#pragma unroll
    for (int c = 0; c < kPanelNB; c++) l[c] = LU[(size_t)i + (size_t)n * (k0 + c0 + min(c, cb - 1))];
#pragma unroll
    for (int c = 0; c < kPanelNB; c++) {
      if (c < cb) {
        const double lr = (double)l[c].x, li = (double)l[c].y;
#pragma unroll
        for (int q = 0; q < kTrsmRhs; q++) {
          const double2 y = xs[q][c0 + c];
          ar[q] += lr * y.x - li * y.y;
          ai[q] += lr * y.y + li * y.x;
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kTrsmRhs; q++)
    if (q < nq) {
      double2 v = X[(size_t)i + (size_t)n * (q0 + q)];
      v.x -= ar[q];
      v.y -= ai[q];
      X[(size_t)i + (size_t)n * (q0 + q)] = v;
    }
}

template <typename R, bool LOWER>
__global__ __launch_bounds__(256) void k_trsm_super(const typename CT<R>::c2* __restrict__ LU,
                                                    const double2* __restrict__ inv, int n, int k0, int kb,
                                                    double2* __restrict__ X, double2* __restrict__ Y, int nrhs,
                                                    const int* __restrict__ skip) {
  if (skip && *skip) return;
  trsm_super_body<R, LOWER>(LU, inv, n, k0, kb, X, Y, nrhs, blockIdx.x);
}

// The same for a batch of systems (as k_tri_inverse_batched, right-hand sides
// at X + x_off[z] with leading dimension nz[z]); systems with no block at k0
// return.
template <typename R, bool LOWER>
__global__ __launch_bounds__(256) void k_trsm_super_batched(const typename CT<R>::c2* __restrict__ LU,
                                                            const size_t* __restrict__ lu_off,
                                                            const int* __restrict__ nz,
                                                            const size_t* __restrict__ inv_off,
                                                            const double2* __restrict__ inv,
                                                            const size_t* __restrict__ x_off, int k0,
                                                            double2* __restrict__ X, double2* __restrict__ Y,
                                                            int nrhs) {
  const int z = blockIdx.z, n = nz[z];
  if (k0 >= n) return;
  trsm_super_body<R, LOWER>(LU + lu_off[z], inv + inv_off[z], n, k0, min(kSuper, n - k0), X + x_off[z],
                            Y + x_off[z], nrhs, blockIdx.x);
}

// The solve kernels for a batch of systems (blockIdx.z = system of b): the
// same bodies on each system's own buffers.
__global__ void k_gather_rows_batch(const SolveBatch b, int n, int nrhs) {
  const SolveMat& m = b.m[blockIdx.z];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (i < n && q < nrhs) m.Xtmp[(size_t)i + (size_t)n * q] = m.X[(size_t)m.perm[i] + (size_t)n * q];
}
__global__ __launch_bounds__(256) void k_tri_inverse_batch(const SolveBatch b, int n) {
  const SolveMat& m = b.m[blockIdx.z];
  tri_inverse_body<double>(m.LU, n, m.inv, blockIdx.x, blockIdx.y);
}
// LOWER: rows updated in Xtmp, solved rows into X; else updated in X, solved into Xtmp
template <bool LOWER>
__global__ __launch_bounds__(256) void k_trsm_super_batch(const SolveBatch b, int n, int k0, int kb, int nrhs) {
  const SolveMat& m = b.m[blockIdx.z];
  trsm_super_body<double, LOWER>(m.LU, m.inv, n, k0, kb, LOWER ? m.Xtmp : m.X, LOWER ? m.X : m.Xtmp, nrhs,
                                 blockIdx.x);
}
__global__ void k_copy_back_batch(const SolveBatch b, size_t count) {
  const SolveMat& m = b.m[blockIdx.z];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    m.X[i] = m.Xtmp[i];
}

__global__ void k_gather_rows(const int* __restrict__ perm, const double2* __restrict__ src, double2* __restrict__ dst,
                              int n, int nrhs, const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (i < n && q < nrhs) dst[(size_t)i + (size_t)n * q] = src[(size_t)perm[i] + (size_t)n * q];
}

__global__ void k_copy_unless(const double2* __restrict__ src, double2* __restrict__ dst, size_t count,
                              const int* __restrict__ skip) {
  if (*skip) return;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// skip (may be null): a device flag; when set, every launch returns at once
// (the refinement loop of the mixed path after convergence).
template <typename R>
static void lu_solve_t(hipStream_t s, const typename CT<R>::c2* LU, int n, const int* perm, double2* X,
                       double2* Xtmp, double2* inv, int nrhs, const int* skip = nullptr) {
  if (nrhs <= 0) return;
  const int groups = (nrhs + kTrsmRhs - 1) / kTrsmRhs;
  // T = P X; forward pass: off-diagonal rows updated in T, solved rows into X;
  // backward pass: off-diagonal rows updated in X, solved rows into T.
  HX_LAUNCH(k_gather_rows, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, perm, X, Xtmp, n, nrhs, skip);
  HX_LAUNCH(k_tri_inverse<R>, dim3((n + kPanelNB - 1) / kPanelNB, 2), dim3(256), 0, s, LU, n, inv, skip);
  for (int k0 = 0; k0 < n; k0 += kSuper) {
    const int kb = n - k0 < kSuper ? n - k0 : kSuper;
    const int rest = n - k0 - kb;
    const int g = rest > 0 ? (rest + 255) / 256 : 1;
    HX_LAUNCH((k_trsm_super<R, true>), dim3(g, groups), dim3(256), 0, s, LU, (const double2*)inv, n, k0, kb, Xtmp,
              X, nrhs, skip);
  }
  for (int k0 = ((n - 1) / kSuper) * kSuper; k0 >= 0; k0 -= kSuper) {
    const int kb = n - k0 < kSuper ? n - k0 : kSuper;
    const int g = k0 > 0 ? (k0 + 255) / 256 : 1;
    HX_LAUNCH((k_trsm_super<R, false>), dim3(g, groups), dim3(256), 0, s, LU, (const double2*)inv, n, k0, kb, X,
              Xtmp, nrhs, skip);
  }
  if (skip)
    HX_LAUNCH(k_copy_unless, dim3(1024), dim3(256), 0, s, Xtmp, X, (size_t)n * nrhs, skip);
  else
    (void)hipMemcpyAsync(X, Xtmp, sizeof(double2) * (size_t)n * nrhs, hipMemcpyDeviceToDevice, s);
}

// Complex-symmetric completion: C(r, c) = C(c, r) for r < c, from the lower
// triangle a syrk left (n x n, column-major).  One 32 x 32 tile of the upper
// triangle per block, read from its mirror tile through LDS so both the
// reads and the writes run down columns.
__global__ __launch_bounds__(256) void k_mirror_lower(double2* __restrict__ C, int n) {
  const int bi = blockIdx.x, bj = blockIdx.y;  // destination tile: rows bi, columns bj
  if (bi > bj) return;
  __shared__ double2 t[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  // source tile: rows bj*32.., columns bi*32..
  for (int k = ty; k < 32; k += 8) {
    const int r = bj * 32 + tx, c = bi * 32 + k;
    if (r < n && c < n) t[k][tx] = C[(size_t)c * n + r];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int r = bi * 32 + tx, c = bj * 32 + k;
    if (r < c && c < n) C[(size_t)c * n + r] = t[tx][k];
  }
}

// ---------------------------------------------------------------------------
// Mixed precision helpers: FP32 copy of A, FP64 residual R = B - A X with
// per-column max norms, and X += D.
// ---------------------------------------------------------------------------
__global__ void k_to_c64(const double2* __restrict__ A, float2* __restrict__ Af, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < count; i += (size_t)gridDim.x * blockDim.x) Af[i] = make_float2((float)A[i].x, (float)A[i].y);
}
__global__ void k_to_c128(const float2* __restrict__ Af, double2* __restrict__ A, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < count; i += (size_t)gridDim.x * blockDim.x) A[i] = make_double2((double)Af[i].x, (double)Af[i].y);
}

// Y = beta B + alpha A X for up to kMvRhs right-hand sides per pass, reading
// every entry of A once (the residual of the mixed path; HBM-bound).  Block
// (rb, chunk, group): 64 rows, kMvChunk columns split over the four waves,
// the wave-uniform X entries come through scalar loads.  Partial sums per
// column chunk go to P and are added in a fixed order by k_matvec_reduce, so
// the result does not depend on scheduling.
constexpr int kMvRhs = 16, kMvChunk = 256;
// right-hand-side group qg of the pass (the kernels below pass blockIdx.z or
// their share of it)
__device__ __forceinline__ void matvec_partial_body(const double2* __restrict__ A, int n,
                                                    const double2* __restrict__ X, int nrhs,
                                                    double2* __restrict__ P, int qg) {
  __shared__ double2 red[3][64][kMvRhs];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane, ic = min(i, n - 1);
  const int q0 = qg * kMvRhs, nq = min(kMvRhs, nrhs - q0);
  const int j0 = blockIdx.y * kMvChunk + w * (kMvChunk / 4);
  const int j1 = min(n, j0 + kMvChunk / 4);
  double ar[kMvRhs], ai[kMvRhs];
#pragma unroll
  for (int q = 0; q < kMvRhs; q++) ar[q] = ai[q] = 0.0;
  for (int j = j0; j < j1; j++) {
    const double2 a = A[(size_t)ic + (size_t)n * j];
#pragma unroll
    for (int q = 0; q < kMvRhs; q++) {  // clamped column: loads stay unconditional
      const double2 x = X[(size_t)j + (size_t)n * (q0 + min(q, nq - 1))];
      ar[q] += a.x * x.x - a.y * x.y;
      ai[q] += a.x * x.y + a.y * x.x;
    }
  }
  if (w > 0)
#pragma unroll
    for (int q = 0; q < kMvRhs; q++) red[w - 1][lane][q] = make_double2(ar[q], ai[q]);
  __syncthreads();
  if (w == 0 && i < n) {
    const size_t base = (size_t)blockIdx.y * n * nrhs;
#pragma unroll
    for (int q = 0; q < kMvRhs; q++)
      if (q < nq) {
        double sr = ar[q], si = ai[q];
        for (int v = 0; v < 3; v++) {
          sr += red[v][lane][q].x;
          si += red[v][lane][q].y;
        }
        P[base + (size_t)i + (size_t)n * (q0 + q)] = make_double2(sr, si);
      }
  }
}
__global__ __launch_bounds__(256) void k_matvec_partial(const double2* __restrict__ A, int n,
                                                        const double2* __restrict__ X, int nrhs,
                                                        double2* __restrict__ P, const int* __restrict__ skip) {
  if (skip && *skip) return;
  matvec_partial_body(A, n, X, nrhs, P, blockIdx.z);
}

__device__ __forceinline__ void matvec_reduce_body(const double2* __restrict__ P, int nchunk, int n, int nrhs,
                                                   const double2* __restrict__ B, double alpha, double beta,
                                                   double2* __restrict__ Y) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x, tot = (size_t)n * nrhs;
  if (e >= tot) return;
  double sr = 0.0, si = 0.0;
  for (int c = 0; c < nchunk; c++) {
    const double2 v = P[(size_t)c * tot + e];
    sr += v.x;
    si += v.y;
  }
  double yr = alpha * sr, yi = alpha * si;
  if (B) {
    yr += beta * B[e].x;
    yi += beta * B[e].y;
  }
  Y[e] = make_double2(yr, yi);
}
__global__ void k_matvec_reduce(const double2* __restrict__ P, int nchunk, int n, int nrhs,
                                const double2* __restrict__ B, double alpha, double beta, double2* __restrict__ Y,
                                const int* __restrict__ skip) {
  if (skip && *skip) return;
  matvec_reduce_body(P, nchunk, n, nrhs, B, alpha, beta, Y);
}

__device__ __forceinline__ void colmax_body(const double2* __restrict__ V, int n, double* __restrict__ norms,
                                            int q) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double m = 0.0;
  for (int i = t; i < n; i += 256) {
    double2 v = V[(size_t)i + (size_t)n * q];
    m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
  }
  red[t] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] = fmax(red[t], red[t + s]);
    __syncthreads();
  }
  if (t == 0) norms[q] = red[0];
}
__global__ void k_colmax(const double2* __restrict__ V, int n, int nrhs, double* __restrict__ norms,
                         const int* __restrict__ skip) {
  if (skip && *skip) return;
  colmax_body(V, n, norms, blockIdx.x);
}

__global__ void k_axpy(double2* __restrict__ X, const double2* __restrict__ D, size_t count,
                       const int* __restrict__ skip) {
  if (skip && *skip) return;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < count; i += (size_t)gridDim.x * blockDim.x) {
    X[i].x += D[i].x;
    X[i].y += D[i].y;
  }
}

// Batched systems stored one after another: rows [r0[z], r0[z+1]) of an
// n x nrhs block X belong to system z and are packed as an nz x nrhs block at
// nrhs * r0[z] (body[i] = the system of row i).
__global__ void k_pack_systems(const double2* __restrict__ X, int n, int nrhs, const int* __restrict__ body,
                               const int* __restrict__ r0, double2* __restrict__ P) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * nrhs) return;
  const int i = (int)(e % n), q = (int)(e / n), z = body[i], a = r0[z], sz = r0[z + 1] - a;
  P[(size_t)nrhs * a + (size_t)(i - a) + (size_t)sz * q] = X[e];
}
__global__ void k_gather_systems(const double2* __restrict__ P, int n, int nrhs, const int* __restrict__ body,
                                 const int* __restrict__ r0, const int* __restrict__ perm, double2* __restrict__ Pt) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * nrhs) return;
  const int i = (int)(e % n), q = (int)(e / n), z = body[i], a = r0[z], sz = r0[z + 1] - a;
  const size_t base = (size_t)nrhs * a + (size_t)sz * q;
  Pt[base + (i - a)] = P[base + perm[i]];  // perm: each system's own row permutation, local indices
}
__global__ void k_unpack_systems(const double2* __restrict__ P, int n, int nrhs, const int* __restrict__ body,
                                 const int* __restrict__ r0, double2* __restrict__ X) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * nrhs) return;
  const int i = (int)(e % n), q = (int)(e / n), z = body[i], a = r0[z], sz = r0[z + 1] - a;
  X[e] = P[(size_t)nrhs * a + (size_t)(i - a) + (size_t)sz * q];
}

// Refinement bookkeeping, after each residual: unless already converged,
// set the converged flag when max over columns of |R|_max / |X|_max <= tol
// (a column with X = 0 compares |R|_max itself), else count the correction
// step that follows.
__device__ __forceinline__ void refine_check_body(const double* __restrict__ norms, int nrhs, double tol,
                                                  int* __restrict__ state) {
  if (state[0]) return;
  double worst = 0.0;
  for (int q = 0; q < nrhs; q++) {
    const double rn = norms[q], xn = norms[nrhs + q];
    worst = fmax(worst, xn > 0.0 ? rn / xn : rn);
  }
  if (worst <= tol)
    state[0] = 1;
  else
    state[1] += 1;  // a correction step follows
}
__global__ void k_refine_check(const double* __restrict__ norms, int nrhs, double tol, int* __restrict__ state) {
  if (threadIdx.x == 0) refine_check_body(norms, nrhs, tol, state);
}

// The refined complex64 solve of a batch of systems (blockIdx.z, or its
// share of blockIdx.z, = system of b): the kernels of refine_solve's steps
// on each system's own buffers, each system skipping what its own state
// skips.  sel picks the vector a step works on: 0 the iterate X, 1 the
// correction D; skip = 1 honours state[0] (converged).
__device__ __forceinline__ double2* rb_vec(const RefineMat& m, int sel) { return sel ? m.D : m.X; }
__device__ __forceinline__ bool rb_skip(const RefineMat& m, int skip) { return skip && m.state[0]; }
__global__ void k_rb_start(const RefineBatch b, size_t count) {  // state reset, B = X
  const RefineMat& m = b.m[blockIdx.z];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    m.state[0] = 0;
    m.state[1] = 0;
  }
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    m.B[i] = m.X[i];
}
__global__ __launch_bounds__(256) void k_rb_tri_inverse(const RefineBatch b, int n) {
  const RefineMat& m = b.m[blockIdx.z];
  tri_inverse_body<float>(m.LU, n, m.inv, blockIdx.x, blockIdx.y);
}
__global__ void k_rb_gather(const RefineBatch b, int n, int nrhs, int sel, int skip) {
  const RefineMat& m = b.m[blockIdx.z];
  if (rb_skip(m, skip)) return;
  const double2* V = rb_vec(m, sel);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (i < n && q < nrhs) m.Xtmp[(size_t)i + (size_t)n * q] = V[(size_t)m.perm[i] + (size_t)n * q];
}
template <bool LOWER>
__global__ __launch_bounds__(256) void k_rb_trsm(const RefineBatch b, int n, int k0, int kb, int nrhs, int sel,
                                                 int skip) {
  const RefineMat& m = b.m[blockIdx.z];
  if (rb_skip(m, skip)) return;
  double2* V = rb_vec(m, sel);
  trsm_super_body<float, LOWER>(m.LU, m.inv, n, k0, kb, LOWER ? m.Xtmp : V, LOWER ? V : m.Xtmp, nrhs, blockIdx.x);
}
__global__ void k_rb_copy_back(const RefineBatch b, size_t count, int sel, int skip) {
  const RefineMat& m = b.m[blockIdx.z];
  if (rb_skip(m, skip)) return;
  double2* V = rb_vec(m, sel);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    V[i] = m.Xtmp[i];
}
__global__ __launch_bounds__(256) void k_rb_matvec_partial(const RefineBatch b, int n, int nrhs, int qgroups) {
  const RefineMat& m = b.m[blockIdx.z / qgroups];
  if (rb_skip(m, 1)) return;
  matvec_partial_body(m.M, n, m.X, nrhs, m.P, blockIdx.z % qgroups);
}
// out: 0 the correction's right-hand side D, 1 the final residual R
__global__ void k_rb_matvec_reduce(const RefineBatch b, int nchunk, int n, int nrhs, int out) {
  const RefineMat& m = b.m[blockIdx.z];
  if (rb_skip(m, 1)) return;
  matvec_reduce_body(m.P, nchunk, n, nrhs, m.B, -1.0, 1.0, out ? m.R : m.D);
}
// which: 0 the residual just formed (out as above) into norms[0, nrhs), 1 X into norms[nrhs, 2 nrhs)
__global__ void k_rb_colmax(const RefineBatch b, int n, int nrhs, int which, int out) {
  const RefineMat& m = b.m[blockIdx.z];
  if (rb_skip(m, 1)) return;
  const double2* V = which ? m.X : (out ? m.R : m.D);
  colmax_body(V, n, m.norms + (which ? nrhs : 0), blockIdx.x);
}
__global__ void k_rb_check(const RefineBatch b, int count, int nrhs, double tol) {
  const int z = threadIdx.x;
  if (z < count) refine_check_body(b.m[z].norms, nrhs, tol, b.m[z].state);
}
__global__ void k_rb_axpy(const RefineBatch b, size_t count) {
  const RefineMat& m = b.m[blockIdx.z];
  if (rb_skip(m, 1)) return;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    m.X[i].x += m.D[i].x;
    m.X[i].y += m.D[i].y;
  }
}
__global__ void k_refine_reset(int* __restrict__ state) {
  state[0] = 0;
  state[1] = 0;
}

namespace launch {
void lu_factor(hipStream_t s, double2* A, int n, int* ipiv, int* perm, int* widx, int* wcnt, int* info,
               double2* diag, const LuAux* aux) {
  LuBatch b{};
  b.m[0] = LuMat{A, ipiv, perm, widx, wcnt, diag, info};
  lu_factor_t<double>(s, b, 1, n, aux);
}
void lu_factor_batch(hipStream_t s, const LuBatch& b, int count, int n, bool c64, const LuAux* aux) {
  if (count < 1 || count > kLuBatchMax) return;
  if (c64)
    lu_factor_t<float>(s, b, count, n, aux);
  else
    lu_factor_t<double>(s, b, count, n, aux);
}
size_t tri_inverse_len(int n) { return (size_t)2 * ((n + kPanelNB - 1) / kPanelNB) * kPanelNB * kPanelNB; }
void lu_solve(hipStream_t s, const double2* LU, int n, const int* perm, double2* X, double2* Xtmp, double2* inv,
              int nrhs) {
  lu_solve_t<double>(s, LU, n, perm, X, Xtmp, inv, nrhs);
}
void lu_solve_batch(hipStream_t s, const SolveBatch& b, int count, int n, int nrhs) {
  if (nrhs <= 0 || count <= 0) return;
  const int groups = (nrhs + kTrsmRhs - 1) / kTrsmRhs;
  HX_LAUNCH(k_gather_rows_batch, dim3((n + 255) / 256, nrhs, count), dim3(256), 0, s, b, n, nrhs);
  HX_LAUNCH(k_tri_inverse_batch, dim3((n + kPanelNB - 1) / kPanelNB, 2, count), dim3(256), 0, s, b, n);
  for (int k0 = 0; k0 < n; k0 += kSuper) {
    const int kb = n - k0 < kSuper ? n - k0 : kSuper;
    const int rest = n - k0 - kb;
    const int g = rest > 0 ? (rest + 255) / 256 : 1;
    HX_LAUNCH(k_trsm_super_batch<true>, dim3(g, groups, count), dim3(256), 0, s, b, n, k0, kb, nrhs);
  }
  for (int k0 = ((n - 1) / kSuper) * kSuper; k0 >= 0; k0 -= kSuper) {
    const int kb = n - k0 < kSuper ? n - k0 : kSuper;
    const int g = k0 > 0 ? (k0 + 255) / 256 : 1;
    HX_LAUNCH(k_trsm_super_batch<false>, dim3(g, groups, count), dim3(256), 0, s, b, n, k0, kb, nrhs);
  }
  HX_LAUNCH(k_copy_back_batch, dim3(256, 1, count), dim3(256), 0, s, b, (size_t)n * nrhs);
}
void mirror_lower(hipStream_t s, double2* C, int n) {
  if (n <= 1) return;
  const unsigned t = (unsigned)((n + 31) / 32);
  HX_LAUNCH(k_mirror_lower, dim3(t, t), dim3(256), 0, s, C, n);
}
void to_c128(hipStream_t s, const float2* Af, double2* A, size_t count) {
  HX_LAUNCH(k_to_c128, dim3(4096), dim3(256), 0, s, Af, A, count);
}
void to_c64(hipStream_t s, const double2* A, float2* Af, size_t count) {
  HX_LAUNCH(k_to_c64, dim3(4096), dim3(256), 0, s, A, Af, count);
}
void lu_factor_c64(hipStream_t s, float2* A, int n, int* ipiv, int* perm, int* widx, int* wcnt, int* info,
                   double2* diag, const LuAux* aux) {
  LuBatch b{};
  b.m[0] = LuMat{A, ipiv, perm, widx, wcnt, diag, info};
  lu_factor_t<float>(s, b, 1, n, aux);
}
void lu_solve_c64(hipStream_t s, const float2* LU, int n, const int* perm, double2* X, double2* Xtmp, double2* inv,
                  int nrhs, const int* skip) {
  lu_solve_t<float>(s, LU, n, perm, X, Xtmp, inv, nrhs, skip);
}
void tri_inverse_systems(hipStream_t s, int nsys, int nmax, const double2* LU, const size_t* lu_off, const int* nz,
                         const size_t* inv_off, double2* inv) {
  HX_LAUNCH(k_tri_inverse_batched<double>, dim3((nmax + kPanelNB - 1) / kPanelNB, 2, nsys), dim3(256), 0, s, LU,
            lu_off, nz, inv_off, inv);
}
void lu_solve_systems(hipStream_t s, int nsys, int nmax, const double2* LU, const size_t* lu_off, const int* nz,
                      const double2* inv, const size_t* inv_off, const size_t* x_off, const int* r0, const int* body,
                      const int* perm, int n, const double2* V, double2* Z, double2* P, double2* Pt, int nrhs) {
  const size_t tot = (size_t)n * nrhs;
  const dim3 g1((unsigned)((tot + 255) / 256));
  const int groups = (nrhs + kTrsmRhs - 1) / kTrsmRhs;
  HX_LAUNCH(k_pack_systems, g1, dim3(256), 0, s, V, n, nrhs, body, r0, P);
  HX_LAUNCH(k_gather_systems, g1, dim3(256), 0, s, P, n, nrhs, body, r0, perm, Pt);
  // forward pass: rows updated in Pt, solved rows into P; backward: updated in P, solved into Pt
  for (int k0 = 0; k0 < nmax; k0 += kSuper) {
    const int rest = nmax - k0 - min(kSuper, nmax - k0);
    const int g = rest > 0 ? (rest + 255) / 256 : 1;
    HX_LAUNCH((k_trsm_super_batched<double, true>), dim3(g, groups, nsys), dim3(256), 0, s, LU, lu_off, nz,
              inv_off, inv, x_off, k0, Pt, P, nrhs);
  }
  for (int k0 = ((nmax - 1) / kSuper) * kSuper; k0 >= 0; k0 -= kSuper) {
    const int g = k0 > 0 ? (k0 + 255) / 256 : 1;
    HX_LAUNCH((k_trsm_super_batched<double, false>), dim3(g, groups, nsys), dim3(256), 0, s, LU, lu_off, nz,
              inv_off, inv, x_off, k0, P, Pt, nrhs);
  }
  HX_LAUNCH(k_unpack_systems, g1, dim3(256), 0, s, Pt, n, nrhs, body, r0, Z);
}
void refine_reset(hipStream_t s, int* state) { HX_LAUNCH(k_refine_reset, dim3(1), dim3(1), 0, s, state); }
void refine_solve_batch(hipStream_t s, const RefineBatch& b, int count, int n, int nrhs, int iters, double tol) {
  if (count <= 0 || nrhs <= 0) return;
  const size_t cnt = (size_t)n * nrhs;
  const int groups = (nrhs + kTrsmRhs - 1) / kTrsmRhs;
  HX_LAUNCH(k_rb_start, dim3(1024, 1, count), dim3(256), 0, s, b, cnt);
  // the complex64 factors' inverted diagonal blocks, once for every solve below
  HX_LAUNCH(k_rb_tri_inverse, dim3((n + kPanelNB - 1) / kPanelNB, 2, count), dim3(256), 0, s, b, n);
  auto solve = [&](int sel, int skip) {
    HX_LAUNCH(k_rb_gather, dim3((n + 255) / 256, nrhs, count), dim3(256), 0, s, b, n, nrhs, sel, skip);
    for (int k0 = 0; k0 < n; k0 += kSuper) {
      const int kb = n - k0 < kSuper ? n - k0 : kSuper;
      const int rest = n - k0 - kb;
      const int g = rest > 0 ? (rest + 255) / 256 : 1;
      HX_LAUNCH(k_rb_trsm<true>, dim3(g, groups, count), dim3(256), 0, s, b, n, k0, kb, nrhs, sel, skip);
    }
    for (int k0 = ((n - 1) / kSuper) * kSuper; k0 >= 0; k0 -= kSuper) {
      const int kb = n - k0 < kSuper ? n - k0 : kSuper;
      const int g = k0 > 0 ? (k0 + 255) / 256 : 1;
      HX_LAUNCH(k_rb_trsm<false>, dim3(g, groups, count), dim3(256), 0, s, b, n, k0, kb, nrhs, sel, skip);
    }
    HX_LAUNCH(k_rb_copy_back, dim3(1024, 1, count), dim3(256), 0, s, b, cnt, sel, skip);
  };
  const int nchunk = (n + kMvChunk - 1) / kMvChunk, qgroups = (nrhs + kMvRhs - 1) / kMvRhs;
  auto residual = [&](int out) {
    HX_LAUNCH(k_rb_matvec_partial, dim3((n + 63) / 64, nchunk, qgroups * count), dim3(256), 0, s, b, n, nrhs,
              qgroups);
    HX_LAUNCH(k_rb_matvec_reduce, dim3((unsigned)((cnt + 255) / 256), 1, count), dim3(256), 0, s, b, nchunk, n, nrhs,
              out);
    HX_LAUNCH(k_rb_colmax, dim3(nrhs, 1, count), dim3(256), 0, s, b, n, nrhs, 0, out);
    HX_LAUNCH(k_rb_colmax, dim3(nrhs, 1, count), dim3(256), 0, s, b, n, nrhs, 1, out);
  };
  solve(0, 0);
  for (int it = 0; it < iters; it++) {
    residual(0);
    HX_LAUNCH(k_rb_check, dim3(1), dim3(64), 0, s, b, count, nrhs, tol);
    solve(1, 1);
    HX_LAUNCH(k_rb_axpy, dim3(1024, 1, count), dim3(256), 0, s, b, cnt);
  }
  residual(1);
}
void refine_check(hipStream_t s, const double* norms, int nrhs, double tol, int* state) {
  HX_LAUNCH(k_refine_check, dim3(1), dim3(64), 0, s, norms, nrhs, tol, state);
}
size_t matvec_workspace(int n, int nrhs) {
  return (size_t)((n + kMvChunk - 1) / kMvChunk) * n * nrhs;
}
void matvec(hipStream_t s, const double2* A, int n, const double2* X, int nrhs, const double2* B, double alpha,
            double beta, double2* Y, double2* P, const int* skip) {
  const int nchunk = (n + kMvChunk - 1) / kMvChunk;
  HX_LAUNCH(k_matvec_partial, dim3((n + 63) / 64, nchunk, (nrhs + kMvRhs - 1) / kMvRhs), dim3(256), 0, s, A, n, X,
            nrhs, P, skip);
  const size_t tot = (size_t)n * nrhs;
  HX_LAUNCH(k_matvec_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, P, nchunk, n, nrhs, B, alpha,
            beta, Y, skip);
}
void residual(hipStream_t s, const double2* A, int n, const double2* X, const double2* B, double2* R, int nrhs,
              double* norms, double2* P, const int* skip) {
  matvec(s, A, n, X, nrhs, B, -1.0, 1.0, R, P, skip);
  HX_LAUNCH(k_colmax, dim3(nrhs), dim3(256), 0, s, R, n, nrhs, norms, skip);
  HX_LAUNCH(k_colmax, dim3(nrhs), dim3(256), 0, s, X, n, nrhs, norms + nrhs, skip);
}
void axpy_update(hipStream_t s, double2* X, const double2* D, size_t count, const int* skip) {
  HX_LAUNCH(k_axpy, dim3(1024), dim3(256), 0, s, X, D, count, skip);
}
}  // namespace launch
}  // namespace hamsx
