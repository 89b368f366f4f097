// hmatrix.hip -- H-matrix operator for the GMRES path, gfx950: ACA
// compression of the admissible blocks of the assembled matrix and the
// H-matrix x block-of-vectors product (HMatrix.f90:556-820 of the reference:
// ADD_LOWRANK_BLOCK, ADD_DENSE_BLOCK, HMATVEC).
//
// Layout (all blocks in cluster order, rows [rs, rs + nr), columns
// [cs, cs + nc); perm maps a cluster position to a panel):
//   * low-rank block: U (nr x rank, column l contiguous) and VH (rank x nc,
//     row l contiguous), A_block ~= U VH;
//   * dense block: D (nr x nc, column-major), a copy of A_block.
// The product runs in three passes: T_b = VH_b x_b for every low-rank block;
// then per leaf cluster of rows, y_i = sum over the blocks that cover row i,
// in a fixed order, of D_b(i, :) x_b or U_b(i, :) T_b; then the scatter back
// to panel order with y = beta B + alpha (H x).  Every output element is
// summed by one thread in a fixed order, so the product is deterministic.
#include <hip/hip_runtime.h>

#include "hamsx_internal.hpp"

namespace hamsx {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double cabs2(double2 a) { return hypot(a.x, a.y); }
__device__ __forceinline__ double2 cdivz(double2 a, double2 b) {  // a / b, Smith
  if (fabs(b.x) >= fabs(b.y)) {
    const double q = b.y / b.x, d = b.x + b.y * q;
    return make_double2((a.x + a.y * q) / d, (a.y - a.x * q) / d);
  }
  const double q = b.x / b.y, d = b.x * q + b.y;
  return make_double2((a.x * q + a.y) / d, (a.y * q - a.x) / d);
}

// block-wide argmax of (v, i): largest v, lowest i on ties; every thread
// gets the result
__device__ __forceinline__ void block_argmax(double& v, int& i, double* sv, int* si) {
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(v, o);
    const int oi = __shfl_xor(i, o);
    if (ov > v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    sv[w] = v;
    si[w] = i;
  }
  __syncthreads();
  v = sv[0];
  i = si[0];
  for (int q = 1; q < 4; q++)
    if (sv[q] > v || (sv[q] == v && si[q] < i)) {
      v = sv[q];
      i = si[q];
    }
}

__device__ __forceinline__ double block_sum(double v, double* sv) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sv[w] = v;
  __syncthreads();
  return (sv[0] + sv[1]) + (sv[2] + sv[3]);
}

// ---------------------------------------------------------------------------
// Partially pivoted ACA of one admissible block per 256-thread workgroup,
// the steps of ADD_LOWRANK_BLOCK (HMatrix.f90:556-750): pivot row 0 first,
// then the unused row with the largest |U(:, k-1)|; residual row, pivot
// column = its largest entry; residual column divided by the pivot; stop
// when |u_k| |v_k| < tol ||U VH||_F, after rank_max steps, on a zero pivot
// or when every row is used.  ||U VH||_F^2 is updated with the exact cross
// terms 2 Re(<u_l, u_k> <v_k, v_l>) of the rank-one sum (the reference's
// update multiplies real parts entry by entry instead, see the note at the
// host caller).  rank[b] receives the rank reached.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_aca(const double2* __restrict__ A, int lda, const int* __restrict__ perm,
                                             const int4* __restrict__ desc, const size_t* __restrict__ uoff,
                                             const size_t* __restrict__ voff, const size_t* __restrict__ roff,
                                             unsigned char* __restrict__ used, double2* __restrict__ U,
                                             double2* __restrict__ VH, int* __restrict__ rank, int rank_max,
                                             double tol) {
  __shared__ double sv[4];
  __shared__ int si[4];
  __shared__ double2 part[4][64];
  const int b = blockIdx.x, t = threadIdx.x;
  const int4 d = desc[b];
  const int rs = d.x, nr = d.y, cs = d.z, nc = d.w;
  const int kmax = min(rank_max, min(nr, nc));
  double2* Ub = U + uoff[b];
  double2* Vb = VH + voff[b];
  unsigned char* ub = used + roff[b];
  for (int i = t; i < nr; i += 256) ub[i] = 0;
  double frob = 0.0;
  int k = 0, pr = 0, rk = 0;
  __syncthreads();
  for (k = 0; k < kmax; k++) {
    if (k > 0) {
      double best = -1.0;
      int bi = 0x7fffffff;
      for (int i = t; i < nr; i += 256)
        if (!ub[i]) {
          const double v = cabs2(Ub[(size_t)(k - 1) * nr + i]);
          if (v > best) {
            best = v;
            bi = i;
          }
        }
      block_argmax(best, bi, sv, si);
      if (best < 0.0) break;  // every row used
      pr = bi;
    }
    __syncthreads();
    if (t == 0) ub[pr] = 1;
    // residual row -> VH(k, :), its largest entry and its squared norm
    const size_t gi = (size_t)perm[rs + pr];
    double rbest = -1.0, vv = 0.0;
    int rj = 0x7fffffff;
    for (int j = t; j < nc; j += 256) {
      double2 v = A[gi + (size_t)lda * perm[cs + j]];
      for (int l = 0; l < k; l++) {
        const double2 p = cmul(Ub[(size_t)l * nr + pr], Vb[(size_t)l * nc + j]);
        v.x -= p.x;
        v.y -= p.y;
      }
      Vb[(size_t)k * nc + j] = v;
      const double a = cabs2(v);
      vv += v.x * v.x + v.y * v.y;
      if (a > rbest) {
        rbest = a;
        rj = j;
      }
    }
    block_argmax(rbest, rj, sv, si);  // its barriers publish the row
    const int pc = rj;
    const double2 piv = Vb[(size_t)k * nc + pc];
    if (cabs2(piv) < 1.0e-30) break;
    vv = block_sum(vv, sv);
    // residual column / pivot -> U(:, k)
    const size_t gj = (size_t)lda * perm[cs + pc];
    double uu = 0.0;
    for (int i = t; i < nr; i += 256) {
      double2 v = A[(size_t)perm[rs + i] + gj];
      for (int l = 0; l < k; l++) {
        const double2 p = cmul(Ub[(size_t)l * nr + i], Vb[(size_t)l * nc + pc]);
        v.x -= p.x;
        v.y -= p.y;
      }
      v = cdivz(v, piv);
      Ub[(size_t)k * nr + i] = v;
      uu += v.x * v.x + v.y * v.y;
    }
    uu = block_sum(uu, sv);  // its barriers publish U(:, k)
    rk = k + 1;
    const double uvn = sqrt(uu) * sqrt(vv);
    frob += uvn * uvn;
    if (k > 0) {
      // cross terms: thread (l, p) sums its quarter of <u_l, u_k> and of <v_k, v_l>
      const int l = t & 63, p = t >> 6;
      double2 su = make_double2(0.0, 0.0), svv = make_double2(0.0, 0.0);
      if (l < k) {
        const int i0 = p * nr / 4, i1 = (p + 1) * nr / 4;
        for (int i = i0; i < i1; i++) {
          const double2 a = Ub[(size_t)l * nr + i], c = Ub[(size_t)k * nr + i];
          su.x += a.x * c.x + a.y * c.y;  // conj(a) c
          su.y += a.x * c.y - a.y * c.x;
        }
        const int j0 = p * nc / 4, j1 = (p + 1) * nc / 4;
        for (int j = j0; j < j1; j++) {
          const double2 a = Vb[(size_t)k * nc + j], c = Vb[(size_t)l * nc + j];
          svv.x += a.x * c.x + a.y * c.y;  // a conj(c)
          svv.y += a.y * c.x - a.x * c.y;
        }
      }
      part[p][l] = su;
      __syncthreads();
      double2 du = make_double2(0.0, 0.0);
      if (p == 0 && l < k)
        for (int q = 0; q < 4; q++) {
          du.x += part[q][l].x;
          du.y += part[q][l].y;
        }
      __syncthreads();
      part[p][l] = svv;
      __syncthreads();
      double cross = 0.0;
      if (p == 0 && l < k) {
        double2 dv = make_double2(0.0, 0.0);
        for (int q = 0; q < 4; q++) {
          dv.x += part[q][l].x;
          dv.y += part[q][l].y;
        }
        cross = du.x * dv.x - du.y * dv.y;  // Re(du dv)
      }
      frob += 2.0 * block_sum(cross, sv);
    }
    if (uvn < tol * sqrt(fmax(frob, 1.0e-30))) break;
  }
  if (t == 0) rank[b] = rk;
}

// dense blocks: D_b = A(perm rows, perm cols), one workgroup per block
__global__ __launch_bounds__(256) void k_dense_gather(const double2* __restrict__ A, int lda,
                                                     const int* __restrict__ perm, const int4* __restrict__ desc,
                                                     const size_t* __restrict__ doff, double2* __restrict__ D) {
  const int4 d = desc[blockIdx.x];
  double2* Db = D + doff[blockIdx.x];
  const size_t tot = (size_t)d.y * d.w;
  for (size_t e = threadIdx.x; e < tot; e += 256) {
    const int i = (int)(e % d.y), j = (int)(e / d.y);
    Db[e] = A[(size_t)perm[d.x + i] + (size_t)lda * perm[d.z + j]];
  }
}

// Xp(k, q) = X(perm[k], q)
__global__ void k_perm_gather(const int* __restrict__ perm, const double2* __restrict__ X, int n, int nrhs,
                              double2* __restrict__ Xp) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * nrhs) return;
  const int k = (int)(e % n), q = (int)(e / n);
  Xp[e] = X[(size_t)perm[k] + (size_t)n * q];
}

// The two product passes stage chunks of kHmJ rows of the right-hand sides
// (kHmQ columns per workgroup, more columns through grid.y) in LDS and
// keep each thread's share of the columns in registers, so every stored
// H entry is read once per kHmQ columns.
constexpr int kHmJ = 16, kHmQ = 128;

__device__ __forceinline__ void stage_rows(double2 (&xs)[kHmJ][kHmQ], const double2* __restrict__ src, size_t ld,
                                           int jn, int nq) {
  for (int e = threadIdx.x; e < kHmJ * kHmQ; e += 256) {
    const int jj = e / kHmQ, qq = e % kHmQ;
    xs[jj][qq] = (jj < jn && qq < nq) ? src[(size_t)jj + ld * qq] : make_double2(0.0, 0.0);
  }
}

// T_b(l, q) = sum_j VH_b(l, j) Xp(cs + j, q), rank_b x nrhs (column-major),
// for low-rank blocks whose ranks are at most 2 KPER, over the columns of one
// work item: item[k] = (block, chunk) sums the chunk's kVxCols columns, so
// a wide block spreads over several workgroups instead of one walking all of
// its columns, and writes the sum at T + out[k] (k_hm_vx_reduce adds a split
// block's chunks).  Workgroup (item, column group of kHmQ): thread (column
// q = t % 128, parity p = t / 128) holds rows l = p + 2k, k < KPER; VH and
// the right-hand-side rows go through LDS in chunks of kHmJ columns.
__device__ __forceinline__ double2 as_d2(double2 v) { return v; }
__device__ __forceinline__ double2 as_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }

// S: the storage of the blocks' factors and dense entries -- double2, or
// float2 (complex64 blocks of the mixed path); arithmetic is FP64 either way.
template <int KPER, typename S>
__global__ __launch_bounds__(256) void k_hm_vx(const int2* __restrict__ item, const size_t* __restrict__ out,
                                               const int4* __restrict__ desc, const int* __restrict__ rank,
                                               const size_t* __restrict__ voff, const S* __restrict__ VH,
                                               const double2* __restrict__ Xp, int n, int nrhs,
                                               double2* __restrict__ T) {
  __shared__ double2 xs[kHmJ][kHmQ];
  __shared__ double2 vs[2 * KPER][kHmJ];
  const int2 it = item[blockIdx.x];
  const int b = it.x;
  const int4 d = desc[b];
  const int rk = rank[b];
  const int jb = it.y * kVxCols, je = min(d.w, jb + kVxCols);  // this item's columns of the block
  const int q0 = blockIdx.y * kHmQ, nq = min(kHmQ, nrhs - q0);
  const int q = threadIdx.x & (kHmQ - 1), p = threadIdx.x >> 7;
  double ar[KPER], ai[KPER];
#pragma unroll
  for (int k = 0; k < KPER; k++) ar[k] = ai[k] = 0.0;
  const S* v = VH + voff[b];
  for (int j0 = jb; j0 < je; j0 += kHmJ) {
    const int jn = min(kHmJ, je - j0);
    __syncthreads();
    stage_rows(xs, Xp + (size_t)d.z + j0 + (size_t)n * q0, (size_t)n, jn, nq);
    for (int e = threadIdx.x; e < 2 * KPER * kHmJ; e += 256) {
      const int l = e / kHmJ, jj = e % kHmJ;
      vs[l][jj] = (l < rk && jj < jn) ? as_d2(v[(size_t)l * d.w + j0 + jj]) : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int jj = 0; jj < jn; jj++) {
      const double2 c = xs[jj][q];
#pragma unroll
      for (int k = 0; k < KPER; k++) {
        const double2 a = vs[p + 2 * k][jj];
        ar[k] += a.x * c.x - a.y * c.y;
        ai[k] += a.x * c.y + a.y * c.x;
      }
    }
  }
  if (q < nq)
#pragma unroll
    for (int k = 0; k < KPER; k++) {
      const int l = p + 2 * k;
      if (l < rk) T[out[blockIdx.x] + (size_t)l + (size_t)rk * (q0 + q)] = make_double2(ar[k], ai[k]);
    }
}

// T of the blocks split into several column chunks: the chunks' partial
// products added in chunk order (a fixed order, so the result does not
// depend on scheduling).  One workgroup per block.
__global__ __launch_bounds__(256) void k_hm_vx_reduce(const int2* __restrict__ red, const size_t* __restrict__ rbase,
                                                      const int* __restrict__ rank, const size_t* __restrict__ toff,
                                                      int nrhs, double2* __restrict__ T) {
  const int2 r = red[blockIdx.x];
  const size_t len = (size_t)rank[r.x] * nrhs;
  const double2* P = T + rbase[blockIdx.x];
  double2* dst = T + toff[r.x];
  for (size_t e = threadIdx.x; e < len; e += 256) {
    double2 acc = P[e];
    for (int c = 1; c < r.y; c++) {
      const double2 v = P[(size_t)c * len + e];
      acc.x += v.x;
      acc.y += v.y;
    }
    dst[e] = acc;
  }
}

// Per segment of a leaf cluster's block list: thread (row lane t / 8, column
// lane t % 8) sums, for its row and columns q0 + lane + 8k, the segment's
// blocks in list order; item >= 0: dense block, item < 0: low-rank block
// ~item.  Workgroup (segment, column group of kHmQ); leaves of more than 32
// rows are done 32 rows at a time.  A leaf of one segment writes its rows of
// Yp; a segment of a longer list writes its partial rows (leading dimension
// the leaf's row count) at Ypart + seg_out, summed by k_hm_leaf_reduce.
template <typename S>
__global__ __launch_bounds__(256) void k_hm_leaf(const int4* __restrict__ seg, const size_t* __restrict__ seg_out,
                                                 const int2* __restrict__ leaves, const int* __restrict__ litem,
                                                 const int4* __restrict__ ddesc,
                                                 const size_t* __restrict__ doff, const S* __restrict__ D,
                                                 const int4* __restrict__ ldesc, const int* __restrict__ rank,
                                                 const size_t* __restrict__ uoff, const size_t* __restrict__ toff,
                                                 const S* __restrict__ U, const double2* __restrict__ T,
                                                 const double2* __restrict__ Xp, int n, int nrhs,
                                                 double2* __restrict__ Yp, double2* __restrict__ Ypart) {
  __shared__ double2 xs[kHmJ][kHmQ];
  const int4 sg = seg[blockIdx.x];
  const int2 lf = leaves[sg.x];
  const int q0 = blockIdx.y * kHmQ, nq = min(kHmQ, nrhs - q0);
  const int lrows = lf.y - lf.x;
  // row r of column q at dst[obase + (r - rbase) + ld q]
  double2* dst = sg.w ? Ypart : Yp;
  const size_t obase = sg.w ? seg_out[blockIdx.x] : 0, ld = sg.w ? (size_t)lrows : (size_t)n;
  const int rbase = sg.w ? lf.x : 0;
  const int rl = threadIdx.x >> 3, ql = threadIdx.x & 7;
  constexpr int kPer = kHmQ / 8;
  for (int r0 = lf.x; r0 < lf.y; r0 += 32) {
    const int row = r0 + rl;
    const bool live = row < lf.y;
    double ar[kPer], ai[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) ar[k] = ai[k] = 0.0;
    for (int it = sg.y; it < sg.z; it++) {
      const int item = litem[it];
      const S* a;           // this row of the block's left factor, column stride nr
      const double2* src;   // right factor rows, leading dimension ld
      size_t ld;
      int nr, kk;
      if (item >= 0) {
        const int4 d = ddesc[item];
        nr = d.y;
        kk = d.w;
        a = D + doff[item] + (live ? row - d.x : 0);
        src = Xp + d.z;
        ld = (size_t)n;
      } else {
        const int lb = ~item;
        const int4 d = ldesc[lb];
        nr = d.y;
        kk = rank[lb];
        a = U + uoff[lb] + (live ? row - d.x : 0);
        src = T + toff[lb];
        ld = (size_t)kk;
      }
      for (int j0 = 0; j0 < kk; j0 += kHmJ) {
        const int jn = min(kHmJ, kk - j0);
        __syncthreads();
        stage_rows(xs, src + j0 + ld * q0, ld, jn, nq);
        __syncthreads();
        if (live)
          for (int jj = 0; jj < jn; jj++) {
            const double2 av = as_d2(a[(size_t)nr * (j0 + jj)]);
#pragma unroll
            for (int k = 0; k < kPer; k++) {
              const double2 c = xs[jj][ql + 8 * k];
              ar[k] += av.x * c.x - av.y * c.y;
              ai[k] += av.x * c.y + av.y * c.x;
            }
          }
      }
    }
    if (live)
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        const int q = ql + 8 * k;
        if (q < nq) dst[obase + (size_t)(row - rbase) + ld * (q0 + q)] = make_double2(ar[k], ai[k]);
      }
  }
}

// Rows of the leaves whose block lists were split into segments: the
// segments' partial rows added in segment order (a fixed order).  One
// workgroup per leaf.
__global__ __launch_bounds__(256) void k_hm_leaf_reduce(const int2* __restrict__ red,
                                                        const size_t* __restrict__ rbase,
                                                        const int2* __restrict__ leaves, int n, int nrhs,
                                                        const double2* __restrict__ Ypart, double2* __restrict__ Yp) {
  const int2 r = red[blockIdx.x];
  const int2 lf = leaves[r.x];
  const int rows = lf.y - lf.x;
  const size_t len = (size_t)rows * nrhs;
  const double2* P = Ypart + rbase[blockIdx.x];
  for (size_t e = threadIdx.x; e < len; e += 256) {
    double2 acc = P[e];
    for (int s = 1; s < r.y; s++) {
      const double2 v = P[(size_t)s * len + e];
      acc.x += v.x;
      acc.y += v.y;
    }
    Yp[(size_t)(lf.x + (int)(e % rows)) + (size_t)n * (e / rows)] = acc;
  }
}

// Y(perm[k], q) = beta B(perm[k], q) + alpha Yp(k, q)   (B may be null)
__global__ void k_perm_scatter(const int* __restrict__ perm, const double2* __restrict__ Yp, int n, int nrhs,
                               const double2* __restrict__ B, double alpha, double beta, double2* __restrict__ Y) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * nrhs) return;
  const int k = (int)(e % n), q = (int)(e / n);
  const size_t o = (size_t)perm[k] + (size_t)n * q;
  double yr = alpha * Yp[e].x, yi = alpha * Yp[e].y;
  if (B) {
    yr += beta * B[o].x;
    yi += beta * B[o].y;
  }
  Y[o] = make_double2(yr, yi);
}

namespace launch {
void hm_aca(hipStream_t s, const double2* A, int lda, const int* perm, int nblk, const int4* desc, const size_t* uoff,
            const size_t* voff, const size_t* roff, unsigned char* used, double2* U, double2* VH, int* rank,
            int rank_max, double tol) {
  if (nblk <= 0) return;
  HX_LAUNCH(k_aca, dim3(nblk), dim3(256), 0, s, A, lda, perm, desc, uoff, voff, roff, used, U, VH, rank, rank_max,
            tol);
}
void hm_dense_gather(hipStream_t s, const double2* A, int lda, const int* perm, int nblk, const int4* desc,
                     const size_t* doff, double2* D) {
  if (nblk <= 0) return;
  HX_LAUNCH(k_dense_gather, dim3(nblk), dim3(256), 0, s, A, lda, perm, desc, doff, D);
}
template <typename S>
static void hm_matvec_t(hipStream_t s, const HmDev& h, const S* U, const S* VH, const S* D, int n,
                        const double2* X, int nrhs, const double2* B, double alpha, double beta, double2* Y,
                        double2* Xp, double2* Yp) {
  const size_t tot = (size_t)n * nrhs;
  const dim3 g1((unsigned)((tot + 255) / 256));
  HX_LAUNCH(k_perm_gather, g1, dim3(256), 0, s, h.perm, X, n, nrhs, Xp);
  const unsigned qg = (unsigned)((nrhs + kHmQ - 1) / kHmQ);
  // low-rank blocks bucketed by rank (<= 8, <= 16, <= 32, <= 64), in work
  // items of kVxCols columns each; then the split blocks' partials summed
  for (int g = 0; g < 4; g++) {
    const int cnt = h.vx_start[g + 1] - h.vx_start[g];
    if (cnt <= 0) continue;
    const dim3 grid(cnt, qg);
    const int2* it = h.vx_item + h.vx_start[g];
    const size_t* out = h.vx_out + h.vx_start[g];
    if (g == 0)
      HX_LAUNCH((k_hm_vx<4, S>), grid, dim3(256), 0, s, it, out, h.ldesc, h.rank, h.voff, VH, Xp, n, nrhs, h.T);
    else if (g == 1)
      HX_LAUNCH((k_hm_vx<8, S>), grid, dim3(256), 0, s, it, out, h.ldesc, h.rank, h.voff, VH, Xp, n, nrhs, h.T);
    else if (g == 2)
      HX_LAUNCH((k_hm_vx<16, S>), grid, dim3(256), 0, s, it, out, h.ldesc, h.rank, h.voff, VH, Xp, n, nrhs, h.T);
    else
      HX_LAUNCH((k_hm_vx<32, S>), grid, dim3(256), 0, s, it, out, h.ldesc, h.rank, h.voff, VH, Xp, n, nrhs, h.T);
  }
  if (h.vx_nred > 0)
    HX_LAUNCH(k_hm_vx_reduce, dim3(h.vx_nred), dim3(256), 0, s, h.vx_red, h.vx_rbase, h.rank, h.toff, nrhs, h.T);
  HX_LAUNCH(k_hm_leaf<S>, dim3(h.nseg, qg), dim3(256), 0, s, h.seg, h.seg_out, h.leaves, h.litem, h.ddesc, h.doff,
            D, h.ldesc, h.rank, h.uoff, h.toff, U, h.T, Xp, n, nrhs, Yp, h.Ypart);
  if (h.leaf_nred > 0)
    HX_LAUNCH(k_hm_leaf_reduce, dim3(h.leaf_nred), dim3(256), 0, s, h.leaf_red, h.leaf_rbase, h.leaves, n, nrhs,
              h.Ypart, Yp);
  HX_LAUNCH(k_perm_scatter, g1, dim3(256), 0, s, h.perm, Yp, n, nrhs, B, alpha, beta, Y);
}
void hm_matvec(hipStream_t s, const HmDev& h, int n, const double2* X, int nrhs, const double2* B, double alpha,
               double beta, double2* Y, double2* Xp, double2* Yp, bool c64) {
  if (c64 && h.U32)
    hm_matvec_t<float2>(s, h, h.U32, h.VH32, h.D32, n, X, nrhs, B, alpha, beta, Y, Xp, Yp);
  else
    hm_matvec_t<double2>(s, h, h.U, h.VH, h.D, n, X, nrhs, B, alpha, beta, Y, Xp, Yp);
}
}  // namespace launch
}  // namespace hamsx
