// krylov.hip -- device side of the restarted GMRES path (ISOLV = 2 of the
// reference, src/IterativeSolvers.f90:190-587), gfx950.
//
// All right-hand sides advance together: column q of every n x nrhs block is
// one independent Krylov space.  Arnoldi vectors V_0..V_m are n x nrhs blocks
// stored one after another.  Orthogonalisation is classical Gram-Schmidt
// applied twice (CGS2), each pass one batched dot kernel and one batched
// update kernel; the small Hessenberg / Givens work runs on the host.  The
// block-diagonal preconditioner packs each body's rows, solves with that
// body's LU factors and unpacks.  Reductions use fixed trees, so results do
// not depend on scheduling.
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
void block_copy(hipStream_t s, const double2* A, int n, int r0, int nb, double2* B) {
  HX_LAUNCH(k_block_copy, g1((size_t)nb * nb), dim3(256), 0, s, A, n, r0, nb, B);
}
}  // namespace launch
}  // namespace hamsx
