// krylov.hip -- device side of the restarted GMRES path (ISOLV = 2 of the
// reference, src/IterativeSolvers.f90:190-587), gfx950.
//
// All right-hand sides advance together: column q of every n x nrhs block is
// one independent Krylov space.  Arnoldi vectors V_0..V_m are n x nrhs blocks
// stored one after another.  Orthogonalisation is classical Gram-Schmidt
// applied twice (CGS2), each pass one batched dot kernel and one batched
// update kernel.  The small Hessenberg / Givens work also runs on the device
// (one thread per right-hand side), so the host enqueues Arnoldi steps
// without waiting for them: it reads the count of unconverged columns a few
// steps behind.  The block-diagonal preconditioner packs each body's rows,
// solves with that body's LU factors and unpacks.  Reductions use fixed
// trees, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include "hamsx_internal.hpp"

namespace hamsx {

// dots[i * nrhs + q] = sum_k conj(V_i(k, q)) W(k, q), i = 0..nv-1 (ZDOT_INLINE
// orientation, IterativeSolvers.f90:82-99); one block per (i, q).
__global__ __launch_bounds__(256) void k_block_dots(const double2* __restrict__ V, int nv, const double2* __restrict__ W,
                                                    int n, int nrhs, double2* __restrict__ dots) {
  __shared__ double rr[256], ri[256];
  const int i = blockIdx.x, q = blockIdx.y, t = threadIdx.x;
  const double2* v = V + ((size_t)i * nrhs + q) * n;
  const double2* w = W + (size_t)q * n;
  double sr = 0.0, si = 0.0;
  for (int k = t; k < n; k += 256) {
    const double2 a = v[k], b = w[k];
    sr += a.x * b.x + a.y * b.y;
    si += a.x * b.y - a.y * b.x;
  }
  rr[t] = sr;
  ri[t] = si;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      rr[t] += rr[t + s];
      ri[t] += ri[t + s];
    }
    __syncthreads();
  }
  if (t == 0) dots[(size_t)i * nrhs + q] = make_double2(rr[0], ri[0]);
}

// W(k, q) -= sum_i h[i * nrhs + q] V_i(k, q), i = 0..nv-1; h accumulates
// into hsum when non-null (the two CGS passes add up).
__global__ void k_block_update(const double2* __restrict__ V, int nv, double2* __restrict__ W, int n, int nrhs,
                               const double2* __restrict__ h, double2* __restrict__ hsum) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x, tot = (size_t)n * nrhs;
  if (e < tot) {
    const int q = (int)(e / n);
    double2 w = W[e];
    for (int i = 0; i < nv; i++) {
      const double2 c = h[(size_t)i * nrhs + q], v = V[(size_t)i * tot + e];
      w.x -= c.x * v.x - c.y * v.y;
      w.y -= c.x * v.y + c.y * v.x;
    }
    W[e] = w;
  }
  if (hsum && blockIdx.x == 0)
    for (int k = threadIdx.x; k < nv * nrhs; k += blockDim.x) {
      hsum[k].x += h[k].x;
      hsum[k].y += h[k].y;
    }
}

// norms[q] = ||W(:, q)||_2 (sum of |re|^2 + |im|^2 as ZNRM2_INLINE, :105-117)
__global__ __launch_bounds__(256) void k_col_norm2(const double2* __restrict__ W, int n, double* __restrict__ norms) {
  __shared__ double red[256];
  const int q = blockIdx.x, t = threadIdx.x;
  double s = 0.0;
  for (int k = t; k < n; k += 256) {
    const double2 v = W[(size_t)q * n + k];
    s += v.x * v.x + v.y * v.y;
  }
  red[t] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) red[t] += red[t + st];
    __syncthreads();
  }
  if (t == 0) norms[q] = sqrt(red[0]);
}

// Y(:, q) = W(:, q) * scale[q]
__global__ void k_col_scale(const double2* __restrict__ W, int n, int nrhs, const double* __restrict__ scale,
                            double2* __restrict__ Y) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * nrhs) return;
  const double f = scale[e / n];
  Y[e] = make_double2(W[e].x * f, W[e].y * f);
}

// U(k, q) = sum_j y[j * nrhs + q] V_j(k, q), j = 0..m-1
__global__ void k_block_combine(const double2* __restrict__ V, int m, int n, int nrhs, const double2* __restrict__ y,
                                double2* __restrict__ U) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x, tot = (size_t)n * nrhs;
  if (e >= tot) return;
  const int q = (int)(e / n);
  double ur = 0.0, ui = 0.0;
  for (int j = 0; j < m; j++) {
    const double2 c = y[(size_t)j * nrhs + q], v = V[(size_t)j * tot + e];
    ur += c.x * v.x - c.y * v.y;
    ui += c.x * v.y + c.y * v.x;
  }
  U[e] = make_double2(ur, ui);
}

// rows [r0, r0 + nb) of X (n x nrhs) <-> a contiguous nb x nrhs block
__global__ void k_rows_pack(const double2* __restrict__ X, int n, int r0, int nb, int nrhs, double2* __restrict__ P) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)nb * nrhs) return;
  const int r = (int)(e % nb), q = (int)(e / nb);
  P[e] = X[(size_t)q * n + r0 + r];
}
__global__ void k_rows_unpack(const double2* __restrict__ P, int n, int r0, int nb, int nrhs, double2* __restrict__ X) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)nb * nrhs) return;
  const int r = (int)(e % nb), q = (int)(e / nb);
  X[(size_t)q * n + r0 + r] = P[e];
}

// diagonal block [r0, r0 + nb)^2 of A (leading dimension n) -> contiguous nb x nb
__global__ void k_block_copy(const double2* __restrict__ A, int n, int r0, int nb, double2* __restrict__ B) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)nb * nb) return;
  const int r = (int)(e % nb), c = (int)(e / nb);
  B[e] = A[(size_t)(r0 + r) + (size_t)n * (r0 + c)];
}

// ---------------------------------------------------------------------------
// Hessenberg / Givens work of GMRES(m) per right-hand side q (the least-squares
// update of IterativeSolvers.f90:387-446).  One 256-thread block loops over
// the columns; GmDev holds the state between launches:
//   H[(q m + j)(m + 1) + i]: column j of q's Hessenberg matrix, rotated;
//   g[j nrhs + q]: the rotated residual vector; cs / sn[j nrhs + q]: rotation j;
//   act[q]: q takes part in this cycle; done[q]: q has converged;
//   mact[q]: columns of q's current cycle; bnorm[q] = ||B(:, q)||.
// out[]: the block's counters for the host (pinned memory): out[0] the
// unconverged columns after a restart test, out[1 + j] the columns still
// active after step j; iters[0] accumulates the longest cycle of each restart.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconjmul(double2 a, double2 b) {  // conj(a) b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double zabs(double2 a) { return hypot(a.x, a.y); }
__device__ __forceinline__ double2 zdiv(double2 a, double2 b) {  // Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const double r = b.x / b.y, d = b.x * r + b.y;
  return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
__device__ __forceinline__ int block_count(int v, int* red) {
  __syncthreads();
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

__global__ __launch_bounds__(256) void k_gm_start(GmDev g, int nrhs, const double* __restrict__ norms) {
  for (int q = threadIdx.x; q < nrhs; q += 256) {
    g.bnorm[q] = norms[q];
    g.done[q] = norms[q] < 1.0e-30 ? 1 : 0;
  }
  if (threadIdx.x == 0) g.iters[0] = 0;
}

// Restart test on beta = ||B - A X||: converged columns leave; the others
// start a cycle with V_0 = R / beta (scale) and g = (beta, 0, ..., 0).
__global__ __launch_bounds__(256) void k_gm_restart(GmDev g, int nrhs, int m, const double* __restrict__ beta,
                                                    double tol, double* __restrict__ scale, int* __restrict__ out,
                                                    double* __restrict__ worst) {
  __shared__ int red[256];
  __shared__ double wred[256];
  int left = 0;
  double w = 0.0;
  for (int q = threadIdx.x; q < nrhs; q += 256) {
    const double b = beta[q], bn = g.bnorm[q];
    if (!g.done[q] && b / bn < tol) g.done[q] = 1;
    if (bn >= 1.0e-30) w = fmax(w, b / bn);
    const int d = g.done[q];
    left += !d;
    scale[q] = d ? 0.0 : 1.0 / b;
    g.act[q] = !d;
    g.mact[q] = 0;
    for (int j = 0; j <= m; j++) g.g[(size_t)j * nrhs + q] = make_double2(0.0, 0.0);
    g.g[q] = make_double2(d ? 0.0 : b, 0.0);
  }
  wred[threadIdx.x] = w;
  left = block_count(left, red);
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) wred[threadIdx.x] = fmax(wred[threadIdx.x], wred[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = left;
    worst[0] = wred[0];
  }
}

// Arnoldi step j: column j of H from the two CGS passes' coefficients d1, d2
// and h(j+1, j) = ||W|| (hn), the earlier rotations applied, rotation j
// formed and applied to g; q stops when |g(j+1)| / ||B(:, q)|| < tol.
// scale[q] = 1 / h(j+1, j) for the columns that go on (0 for the others).
__global__ __launch_bounds__(256) void k_gm_step(GmDev g, int nrhs, int m, int j, const double2* __restrict__ d1,
                                                 const double2* __restrict__ d2, const double* __restrict__ hn,
                                                 double tol, double* __restrict__ scale, int* __restrict__ out) {
  __shared__ int red[256];
  int live = 0;
  for (int q = threadIdx.x; q < nrhs; q += 256) {
    scale[q] = 0.0;
    if (!g.act[q]) continue;
    double2* h = g.H + ((size_t)q * m + j) * (m + 1);
    for (int i = 0; i <= j; i++) {
      const double2 a = d1[(size_t)i * nrhs + q], b = d2[(size_t)i * nrhs + q];
      h[i] = make_double2(a.x + b.x, a.y + b.y);
    }
    h[j + 1] = make_double2(hn[q], 0.0);
    for (int i = 0; i < j; i++) {
      const double ci = g.cs[(size_t)i * nrhs + q];
      const double2 si = g.sn[(size_t)i * nrhs + q];
      const double2 hi = h[i], hi1 = h[i + 1], shi1 = zmul(si, hi1), cshi = zconjmul(si, hi);
      h[i] = make_double2(ci * hi.x + shi1.x, ci * hi.y + shi1.y);
      h[i + 1] = make_double2(ci * hi1.x - cshi.x, ci * hi1.y - cshi.y);
    }
    // rotation j: cs = |a| / d, sn = (a / |a|) conj(b) / d, d = sqrt(|a|^2 + |b|^2)
    const double2 a = h[j], b = h[j + 1];
    const double ab = zabs(a), bb = zabs(b);
    double cj;
    double2 sj;
    if (bb == 0.0) {
      cj = 1.0;
      sj = make_double2(0.0, 0.0);
    } else if (ab == 0.0) {
      cj = 0.0;
      sj = make_double2(1.0, 0.0);
    } else {
      const double dd = sqrt(ab * ab + bb * bb);
      cj = ab / dd;
      const double2 u = make_double2(a.x / ab, a.y / ab), t = zmul(u, make_double2(b.x, -b.y));
      sj = make_double2(t.x / dd, t.y / dd);
    }
    g.cs[(size_t)j * nrhs + q] = cj;
    g.sn[(size_t)j * nrhs + q] = sj;
    const double2 sb = zmul(sj, b);
    h[j] = make_double2(cj * a.x + sb.x, cj * a.y + sb.y);
    h[j + 1] = make_double2(0.0, 0.0);
    const double2 gj = g.g[(size_t)j * nrhs + q], gj1 = g.g[(size_t)(j + 1) * nrhs + q];
    const double2 sg1 = zmul(sj, gj1), csg = zconjmul(sj, gj);
    g.g[(size_t)j * nrhs + q] = make_double2(cj * gj.x + sg1.x, cj * gj.y + sg1.y);
    const double2 ng1 = make_double2(cj * gj1.x - csg.x, cj * gj1.y - csg.y);
    g.g[(size_t)(j + 1) * nrhs + q] = ng1;
    g.mact[q] = j + 1;
    if (zabs(ng1) / g.bnorm[q] < tol) {
      g.act[q] = 0;  // converged in this cycle: its later columns are zero
    } else {
      live++;
      scale[q] = hn[q] > 1.0e-30 ? 1.0 / hn[q] : 0.0;
    }
  }
  live = block_count(live, red);
  if (threadIdx.x == 0) out[1 + j] = live;
}

// End of a cycle: y = H^-1 g per right-hand side (back substitution over its
// mact columns), y[i nrhs + q] for i < m, zero past mact; iters[0] += the
// longest cycle.
__global__ __launch_bounds__(256) void k_gm_solve(GmDev g, int nrhs, int m, double2* __restrict__ y) {
  __shared__ int red[256];
  int longest = 0;
  for (int q = threadIdx.x; q < nrhs; q += 256) {
    const int mq = g.mact[q];
    longest = max(longest, mq);
    for (int i = 0; i < m; i++) y[(size_t)i * nrhs + q] = i < mq ? g.g[(size_t)i * nrhs + q] : make_double2(0.0, 0.0);
    for (int i = mq - 1; i >= 0; i--) {
      const double2* hc = g.H + ((size_t)q * m + i) * (m + 1);  // column i
      const double2 hii = hc[i];
      double2 yi = y[(size_t)i * nrhs + q];
      yi = zabs(hii) < 1.0e-30 ? make_double2(0.0, 0.0) : zdiv(yi, hii);
      y[(size_t)i * nrhs + q] = yi;
      for (int k = 0; k < i; k++) {
        const double2 t = zmul(hc[k], yi), yk = y[(size_t)k * nrhs + q];
        y[(size_t)k * nrhs + q] = make_double2(yk.x - t.x, yk.y - t.y);
      }
    }
  }
  __syncthreads();
  red[threadIdx.x] = longest;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) g.iters[0] += red[0];
}

namespace launch {
static dim3 g1(size_t count) { return dim3((unsigned)((count + 255) / 256)); }

void block_dots(hipStream_t s, const double2* V, int nv, const double2* W, int n, int nrhs, double2* dots) {
  HX_LAUNCH(k_block_dots, dim3(nv, nrhs), dim3(256), 0, s, V, nv, W, n, nrhs, dots);
}
void block_update(hipStream_t s, const double2* V, int nv, double2* W, int n, int nrhs, const double2* h,
                  double2* hsum) {
  HX_LAUNCH(k_block_update, g1((size_t)n * nrhs), dim3(256), 0, s, V, nv, W, n, nrhs, h, hsum);
}
void col_norm2(hipStream_t s, const double2* W, int n, int nrhs, double* norms) {
  HX_LAUNCH(k_col_norm2, dim3(nrhs), dim3(256), 0, s, W, n, norms);
}
void col_scale(hipStream_t s, const double2* W, int n, int nrhs, const double* scale, double2* Y) {
  HX_LAUNCH(k_col_scale, g1((size_t)n * nrhs), dim3(256), 0, s, W, n, nrhs, scale, Y);
}
void block_combine(hipStream_t s, const double2* V, int m, int n, int nrhs, const double2* y, double2* U) {
  HX_LAUNCH(k_block_combine, g1((size_t)n * nrhs), dim3(256), 0, s, V, m, n, nrhs, y, U);
}
void rows_pack(hipStream_t s, const double2* X, int n, int r0, int nb, int nrhs, double2* P) {
  HX_LAUNCH(k_rows_pack, g1((size_t)nb * nrhs), dim3(256), 0, s, X, n, r0, nb, nrhs, P);
}
void rows_unpack(hipStream_t s, const double2* P, int n, int r0, int nb, int nrhs, double2* X) {
  HX_LAUNCH(k_rows_unpack, g1((size_t)nb * nrhs), dim3(256), 0, s, P, n, r0, nb, nrhs, X);
}
void gm_start(hipStream_t s, const GmDev& g, int nrhs, const double* norms) {
  HX_LAUNCH(k_gm_start, dim3(1), dim3(256), 0, s, g, nrhs, norms);
}
void gm_restart(hipStream_t s, const GmDev& g, int nrhs, int m, const double* beta, double tol, double* scale,
                int* out, double* worst) {
  HX_LAUNCH(k_gm_restart, dim3(1), dim3(256), 0, s, g, nrhs, m, beta, tol, scale, out, worst);
}
void gm_step(hipStream_t s, const GmDev& g, int nrhs, int m, int j, const double2* d1, const double2* d2,
             const double* hn, double tol, double* scale, int* out) {
  HX_LAUNCH(k_gm_step, dim3(1), dim3(256), 0, s, g, nrhs, m, j, d1, d2, hn, tol, scale, out);
}
void gm_solve(hipStream_t s, const GmDev& g, int nrhs, int m, double2* y) {
  HX_LAUNCH(k_gm_solve, dim3(1), dim3(256), 0, s, g, nrhs, m, y);
}
void block_copy(hipStream_t s, const double2* A, int n, int r0, int nb, double2* B) {
  HX_LAUNCH(k_block_copy, g1((size_t)nb * nb), dim3(256), 0, s, A, n, r0, nb, B);
}
}  // namespace launch
}  // namespace hamsx
