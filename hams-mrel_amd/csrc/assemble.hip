// assemble.hip -- influence-matrix assembly, right-hand sides and force
// reductions for one frequency on gfx950.
//
// Assembly is one thread per (field i, source j) pair: a 256-thread block is
// four wave64s; each wave owns one source panel j (wave-uniform, so the
// source geometry and Newman frame are scalar loads) and 64 consecutive
// field panels i, so each wave writes 64 consecutive complex128 entries of
// column j of A and of S (column-major, coalesced 1 KiB stores).  The pair
// loop of CALGREEN_MULTI and the entry formulas of BODINT_MULTI_LEFT /
// RBC_MULTI_RIGHT are fused: no Green-function intermediates reach HBM.
#include "green_dev.hpp"
#include "hamsx_internal.hpp"

namespace hamsx {

// FLAG = DIST <= 50*PNSZ(J) (CalGreenFuncMulti.f90:134-139), formed without
// FMA contraction so the decision matches the FP64 reference bit for bit.
__device__ __forceinline__ int near_flag(double xi, double yi, double zi, double xj, double yj, double zj, double pn) {
#pragma clang fp contract(off)
  double dx = xi - xj, dy = yi - yj, dz = zi - zj;
  double dist = sqrt(dx * dx + dy * dy + dz * dz);
  return dist <= 50.0 * pn ? 1 : 0;
}

// n: the table stride (field points)
__device__ __forceinline__ FreqParams freq_params(const FreqDev& f, int n, int js, int jf) {
  const double* t = f.tab;
  const size_t mn = (size_t)kModes * n;
  const double* p = t + kTabHead;
  return FreqParams{f.v,     f.wk,       f.w1,           f.h,    f.wvn,          t,  t + kModes, t + 2 * kModes,
                    t[3 * kModes], p, p + mn, p + 2 * mn, p + 3 * mn, p + 4 * mn, p + 4 * mn + n, n, js, jf,
                    f.far32,   f.limes_ieee};
}

// Mode tables: one thread per (panel, mode); panel 0's threads also write
// the per-mode scalars, and thread (0, 0) N_0.
__global__ __launch_bounds__(256) void k_mode_tables(MeshDev m, FreqDev f) {
#pragma clang fp contract(off)
  f = freq_now(f);
  const int panel = blockIdx.x * 256 + threadIdx.x, mode = blockIdx.y, n = m.nf;
  if (panel >= n) return;
  double* t = f.tab;
  const size_t mn = (size_t)kModes * n;
  double* p = t + kTabHead;
  double cz, sz, ct, st;
  mode_panel(mode, f.wvn, f.h, m.fz[panel], cz, sz, ct, st);
  const size_t o = (size_t)mode * n + panel;
  p[o] = cz;
  p[mn + o] = sz;
  p[2 * mn + o] = ct;
  p[3 * mn + o] = st;
  if (mode == 0) {
    const double z = m.fz[panel];
    p[4 * mn + panel] = cosh(f.wk * (z + f.h));
    p[4 * mn + n + panel] = sinh(f.wk * (z + f.h));
  }
  if (panel == 0) {
    double pm, ac, ut;
    mode_scalars(mode, f.wvn, f.v, f.h, pm, ac, ut);
    t[mode] = pm;
    t[kModes + mode] = ac;
    t[2 * kModes + mode] = ut;
    if (mode == 0) t[3 * kModes] = wave_nm0(f.wk, f.h);
  }
}

// A(i,j) and S(i,j) from the Green function g and, when FLAG = 1, the
// Newman integrals of source panel j (BODINT_MULTI_LEFT / RBC_MULTI_RIGHT).
__device__ __forceinline__ void store_entry(const MeshDev& m, int i, int j, int flag, double sij, const double dij[3],
                                            const cd g[4], double2* __restrict__ A, double2* __restrict__ S) {
  const int n = m.n, ld = m.nf - m.row0;
  const double ear = m.area[j], n1 = m.nrm[j], n2 = m.nrm[n + j], n3 = m.nrm[2 * n + j];
  cd t, s;
  if (flag) {
    t = (mk(dij[0]) + g[1] * ear) * n1 + (mk(dij[1]) + g[2] * ear) * n2 + (mk(dij[2]) + g[3] * ear) * n3;
    s = mk(sij) + g[0] * ear;
  } else {
    t = (g[1] * n1 + g[2] * n2 + g[3] * n3) * ear;
    s = g[0] * ear;
  }
  // AMAT(IEL,IEL) = 2*PI before the sum (AssbMatxMulti.f90:72-74); lid rows have
  // none (AssbMatx_irrMulti.f90:87-89)
  if (i == j) t.r = 2.0 * kPi + t.r;
  const size_t o = (size_t)(i - m.row0) + (size_t)ld * j;
  A[o] = make_double2(t.r, t.i);
  S[o] = make_double2(s.r, s.i);
}

// One pair: the Green function (deep water, or finite depth in region REG),
// the Newman integrals when FLAG = 1, and the entries A(i,j), S(i,j).
template <bool DEEP, int REG>
__device__ __forceinline__ void pair_entry(const MeshDev& m, const FreqDev& f, int i, int j, double2* __restrict__ A,
                                           double2* __restrict__ S, int* flags) {
  const int nf = m.nf;
  const double xj = m.cx[j], yj = m.cy[j], zj = m.cz[j];
  const double xi = m.fx[i], yi = m.fy[i], zi = m.fz[i];
  const int flag = near_flag(m.gx[i], m.gy[i], m.gz[i], xj, yj, zj, m.pnsz[j]);
  double sij = 0.0, dij[3] = {0.0, 0.0, 0.0};
  if (flag) d_newman(m.frame + (size_t)kFrameLen * j, m.ncn[j], xi, yi, zi, sij, dij);
  PairGeom p = pair_geom(xj, xi, yj, yi, zj, zi);
  cd g[4];
  int ovf = 0;
  if (DEEP) {
    d_infgreen(p, f.v, flag, g);
  } else {
    const FreqParams fp = freq_params(f, nf, j, i);
    d_fingreen_r<REG>(p, fp, flag, g, &ovf);
  }
  if (ovf) atomicOr(flags, 1);
  store_entry(m, i, j, flag, sij, dij, g, A, S);
}

// Both orientations of hull pair {a, b} in region REG (PIDCOCK, EIGENE or
// EIGEN): (field a, source b) and (field b, source a), the Bessel functions
// of their common R evaluated once.  Each entry equals what pair_entry
// gives for it.  Hull field point k is hull centroid k.
template <int REG>
__device__ __forceinline__ void pair_entry2(const MeshDev& m, const FreqDev& f, int a, int b,
                                            double2* __restrict__ A, double2* __restrict__ S) {
  const int nf = m.nf;
  const double xa = m.fx[a], ya = m.fy[a], za = m.fz[a], xb = m.fx[b], yb = m.fy[b], zb = m.fz[b];
  const double sxa = m.cx[a], sya = m.cy[a], sza = m.cz[a], sxb = m.cx[b], syb = m.cy[b], szb = m.cz[b];
  const int flagA = near_flag(xa, ya, za, sxb, syb, szb, m.pnsz[b]);
  const int flagB = near_flag(xb, yb, zb, sxa, sya, sza, m.pnsz[a]);
  double sA = 0.0, dA[3] = {0.0, 0.0, 0.0}, sB = 0.0, dB[3] = {0.0, 0.0, 0.0};
  if (flagA) d_newman(m.frame + (size_t)kFrameLen * b, m.ncn[b], xa, ya, za, sA, dA);
  if (flagB) d_newman(m.frame + (size_t)kFrameLen * a, m.ncn[a], xb, yb, zb, sB, dB);
  const PairGeom pA = pair_geom(sxb, xa, syb, ya, szb, za), pB = pair_geom(sxa, xb, sya, yb, sza, zb);
  const FreqParams fpA = freq_params(f, nf, b, a), fpB = freq_params(f, nf, a, b);
  cd gA[4], gB[4];
  d_fingreen2<REG>(pA, pB, fpA, fpB, flagA, flagB, gA, gB);
  store_entry(m, a, b, flagA, sA, dA, gA, A, S);
  store_entry(m, b, a, flagB, sB, dB, gB, A, S);
}

// Grid form: a 256-thread block is four wave64s; each wave owns one source
// panel j (wave-uniform, so the source geometry and Newman frame are scalar
// loads) and 64 consecutive field panels i, writing 64 consecutive complex128
// entries of column j of A and of S.  Used for deep water, whose Green
// function has no long region-dependent series.
template <bool DEEP>
__global__ __launch_bounds__(256) void k_assemble(MeshDev m, FreqDev f, double2* __restrict__ A,
                                                  double2* __restrict__ S, int* flags) {
  f = freq_now(f);
  const int lane = threadIdx.x & 63;
  const int j = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + (threadIdx.x >> 6));
  const int i = m.row0 + blockIdx.x * 64 + lane;
  if (j >= m.n || i >= m.nf) return;
  pair_entry<DEEP, kRegAny>(m, f, i, j, A, S, flags);
}

// ---------------------------------------------------------------------------
// Finite depth, binned form.  The pairs are counting-sorted once per
// (mesh, h) by the keys of hamsx_internal.hpp into a list of packed
// (j << 16 | i) indices -- for a hull pair entry i = a < j = b -- each bin
// padded to a multiple of 64 with kNoPair, so every wave runs one region
// with one series length and one FLAG pattern.  Per region one kernel walks
// the single entries and one the hull-pair entries.
// ---------------------------------------------------------------------------
constexpr unsigned kNoPair = 0xFFFFFFFFu;

// key of grid position (field i, source j), or -1 when the pair is listed
// under its hull partner (j, i)
__device__ __forceinline__ int pair_key(const MeshDev& m, double h, int i, int j) {
  const double xj = m.cx[j], yj = m.cy[j], zj = m.cz[j];
  const double xi = m.fx[i], yi = m.fy[i], zi = m.fz[i];
  const int flag = near_flag(m.gx[i], m.gy[i], m.gz[i], xj, yj, zj, m.pnsz[j]);
  PairGeom p = pair_geom(xj, xi, yj, yi, zj, zi);
  const int sub = fin_subkey(p.rr, h);
  if (sub >= kSubPidcock && i < m.n && i != j) {
    if (i > j) return -1;
    const int flag_ba = near_flag(m.fx[j], m.fy[j], m.fz[j], m.cx[i], m.cy[i], m.cz[i], m.pnsz[i]);
    return kPairKey0 + 4 * (sub - kSubPidcock) + 2 * flag + flag_ba;
  }
  return 2 * sub + flag;
}

__global__ __launch_bounds__(256) void k_bin_count(MeshDev m, double h, int* __restrict__ count) {
  __shared__ int hist[kFinBins];
  for (int k = threadIdx.x; k < kFinBins; k += 256) hist[k] = 0;
  __syncthreads();
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6), i = m.row0 + blockIdx.x * 64 + (threadIdx.x & 63);
  if (i < m.nf && j < m.n) {
    const int key = pair_key(m, h, i, j);
    if (key >= 0) atomicAdd(&hist[key], 1);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kFinBins; k += 256)
    if (hist[k]) atomicAdd(&count[k], hist[k]);
}

// cursor[] starts at each bin's (padded) offset and ends at offset + count.
__global__ __launch_bounds__(256) void k_bin_scatter(MeshDev m, double h, int* __restrict__ cursor,
                                                     unsigned* __restrict__ list) {
  __shared__ int hist[kFinBins], base[kFinBins];
  for (int k = threadIdx.x; k < kFinBins; k += 256) hist[k] = 0;
  __syncthreads();
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6), i = m.row0 + blockIdx.x * 64 + (threadIdx.x & 63);
  int key = -1, rank = 0;
  if (i < m.nf && j < m.n) {
    key = pair_key(m, h, i, j);
    if (key >= 0) rank = atomicAdd(&hist[key], 1);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kFinBins; k += 256)
    if (hist[k]) base[k] = atomicAdd(&cursor[k], hist[k]);
  __syncthreads();
  if (key >= 0) list[base[key] + rank] = ((unsigned)j << 16) | (unsigned)i;
}

// EIGENE's register-resident epsilon tables take ~200 VGPRs: asking for two
// waves per SIMD keeps the kernel within 256 registers (a few spills)
// instead of one wave per SIMD.
template <int REG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(REG == kRegEigene || REG == kRegPidcock ? 2 : 1))) void k_assemble_list(MeshDev m, FreqDev f, const unsigned* __restrict__ list,
                                                       int count, double2* __restrict__ A, double2* __restrict__ S,
                                                       int* flags) {
  f = freq_now(f);
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= count) return;
  const unsigned e = list[t];
  if (e == kNoPair) return;
  pair_entry<false, REG>(m, f, (int)(e & 0xFFFFu), (int)(e >> 16), A, S, flags);
}

// Hull-pair entries: both orientations per thread.
template <int REG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(REG == kRegEigene || REG == kRegPidcock ? 2 : 1))) void k_assemble_pairs(MeshDev m, FreqDev f, const unsigned* __restrict__ list,
                                                        int count, double2* __restrict__ A,
                                                        double2* __restrict__ S) {
  f = freq_now(f);
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= count) return;
  const unsigned e = list[t];
  if (e == kNoPair) return;
  pair_entry2<REG>(m, f, (int)(e & 0xFFFFu), (int)(e >> 16), A, S);
}

// ---------------------------------------------------------------------------
// Incident wave (PatclVelct.f90:42-123).  DINP returns the velocity
// components divided by CI; VINP the potential divided by CI.
// ---------------------------------------------------------------------------
__device__ inline void d_wave_consts(const FreqDev& f, double& w0, double& wk0) {
  if (fabs(f.v) < HX_F(1.0e-8)) {
    w0 = HX_F(1.0e-20);
    wk0 = HX_F(1.0e-20);
  } else if (fabs(f.v + 1.0) < HX_F(1.0e-8)) {
    w0 = HX_F(1.0e20);
    wk0 = HX_F(1.0e20);
  } else {
    w0 = f.w1;
    wk0 = f.wk;
  }
}

__device__ inline void d_dinp(double x, double y, double z, const FreqDev& f, const WaveDev& w, double beta, cd out[3]) {
  double w0, wk0;
  d_wave_consts(f, w0, wk0);
  double sb, cb;
  sincos(beta, &sb, &cb);
  double wkx = wk0 * ((x - w.xw0) * cb + (y - w.xw1) * sb);
  double se, ce;
  sincos(wkx, &se, &ce);
  double th, tv;
  if (f.h < 0.0) {
    th = exp(wk0 * z);
    tv = th;
  } else {
    double sh = sinh(wk0 * f.h);
    th = cosh(wk0 * (z + f.h)) / sh;
    tv = sinh(wk0 * (z + f.h)) / sh;
  }
  double ah = w0 * w.amp * th, av = w0 * w.amp * tv;
  // (a e^{i wkx}) / i = a (sin, -cos)
  out[0] = cd{ah * se * cb, -(ah * ce * cb)};
  out[1] = cd{ah * se * sb, -(ah * ce * sb)};
  // (-i a e^{i wkx}) / i = -a e^{i wkx}
  out[2] = cd{-(av * ce), -(av * se)};
}

__device__ inline cd d_vinp(double x, double y, double z, const FreqDev& f, const WaveDev& w, double beta) {
  double w0, wk0;
  d_wave_consts(f, w0, wk0);
  double sb, cb;
  sincos(beta, &sb, &cb);
  double wkx = wk0 * ((x - w.xw0) * cb + (y - w.xw1) * sb);
  double se, ce;
  sincos(wkx, &se, &ce);
  if (f.h < 0.0 || f.h > 500.0) {
    // (-i AMP G / W0 e^{wk0 z} e^{i wkx}) / i
    double a = -(w.amp * kG / w0 * exp(wk0 * z));
    return cd{a * ce, a * se};
  }
  double dum = -w.amp * w0 / (wk0 * sinh(wk0 * f.h));
  double a = dum * cosh(wk0 * (z + f.h));
  return cd{a * ce, a * se};  // (i a e^{i wkx}) / i
}

// dum2(j, beta) = grad(phi_I) . n at panel j (DBC_MULTI_RIGHT, BodyIntgrMulti.f90:150-163)
__global__ void k_dum2(MeshDev m, FreqDev f, WaveDev w, double2* dum2) {
  f = freq_now(f);
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (j >= m.n) return;
  cd d[3];
  d_dinp(m.cx[j], m.cy[j], m.cz[j], f, w, w.beta[b], d);
  cd v = d[0] * m.nrm[j] + d[1] * m.nrm[m.n + j] + d[2] * m.nrm[2 * m.n + j];
  dum2[(size_t)b * m.n + j] = make_double2(v.r, v.i);
}

// Partial products of S with the generator columns over one chunk of source
// panels (all in one body): 6 radiation columns of that body (n_1..n_6,
// ASSB_RBC_MULTI) and nbeta diffraction columns (-dum2, ASSB_DBC_MULTI).
constexpr int kBetaGroup = 8;
__global__ __launch_bounds__(64) void k_rhs_partial(int n, int nf, const double* __restrict__ nrm,
                                                    const double2* __restrict__ S, const double2* __restrict__ dum2,
                                                    int nbeta, const int* chunk_j0, const int* chunk_j1,
                                                    double2* __restrict__ partial) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int c = blockIdx.y;
  const int g0 = blockIdx.z * kBetaGroup;
  if (i >= nf) return;
  const int j0 = chunk_j0[c], j1 = chunk_j1[c];
  const int nslot = 6 + nbeta;
  cd rad[6], dif[kBetaGroup];
  for (int k = 0; k < 6; k++) rad[k] = mk(0.0);
  for (int k = 0; k < kBetaGroup; k++) dif[k] = mk(0.0);
  const int ng = min(kBetaGroup, nbeta - g0);
  for (int j = j0; j < j1; j++) {
    double2 sv = S[(size_t)i + (size_t)nf * j];
    cd s{sv.x, sv.y};
    if (blockIdx.z == 0)
      for (int k = 0; k < 6; k++) rad[k] = rad[k] + s * nrm[(size_t)k * n + j];
    for (int b = 0; b < ng; b++) {
      double2 dv = dum2[(size_t)(g0 + b) * n + j];
      dif[b] = dif[b] - s * cd{dv.x, dv.y};
    }
  }
  double2* out = partial + (size_t)c * nslot * nf;
  if (blockIdx.z == 0)
    for (int k = 0; k < 6; k++) out[(size_t)k * nf + i] = make_double2(rad[k].r, rad[k].i);
  for (int b = 0; b < ng; b++) out[(size_t)(6 + g0 + b) * nf + i] = make_double2(dif[b].r, dif[b].i);
}

// Sum the chunk partials in chunk order: X(:, 6b+m) over the chunks of body b,
// X(:, 6nb+beta) over all chunks (n = rows = leading dimension).
__global__ void k_rhs_reduce(int n, int nbody, int nbeta, int nchunk, const int* chunk_body,
                             const double2* __restrict__ partial, double2* __restrict__ X) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int col = blockIdx.y;
  if (i >= n) return;
  const int nslot = 6 + nbeta;
  double re = 0.0, im = 0.0;
  if (col < 6 * nbody) {
    int b = col / 6, slot = col % 6;
    for (int c = 0; c < nchunk; c++)
      if (chunk_body[c] == b) {
        double2 v = partial[((size_t)c * nslot + slot) * n + i];
        re += v.x;
        im += v.y;
      }
  } else {
    int slot = 6 + (col - 6 * nbody);
    for (int c = 0; c < nchunk; c++) {
      double2 v = partial[((size_t)c * nslot + slot) * n + i];
      re += v.x;
      im += v.y;
    }
  }
  X[(size_t)col * n + i] = make_double2(re, im);
}

// ---------------------------------------------------------------------------
// Force reductions (RFORCE_MULTI / EFORCE_MULTI, PotentWavForceMulti.f90:45-226).
// One block per (body M, potential column r); 256 threads reduce over the
// body's panels in a fixed order.
//   r < 6nb : AMAS(6M+k, r) = sum Im(i rho phi DS n_k),
//             BDMP(6M+k, r) = -sum Re(i rho w phi DS n_k)
//   r >= 6nb: EXFC(6M+k, beta) = i w rho sum (phi_D [+ phi_I]) DS n_k
// ---------------------------------------------------------------------------
// With a symmetry plane (mirror 0: x -> -x, 1: y -> -y) the mirror half's
// potentials Xm add their terms with the RX / RY signs of the normals and
// VINP at the mirrored centroids (the IP = 2 terms of RFORCE/EFORCE_MULTI).
__global__ __launch_bounds__(256) void k_forces(MeshDev m, FreqDev f, WaveDev w, int nbody, const int* body_start,
                                                const double2* __restrict__ X, const double2* __restrict__ Xm,
                                                int mirror, double* __restrict__ out) {
  f = freq_now(f);
  __shared__ double red[12][256];
  const int bm = blockIdx.x, r = blockIdx.y, t = threadIdx.x;
  const int n = m.n, nd = 6 * nbody;
  const bool rad = r < nd;
  double acc[12];
  for (int k = 0; k < 12; k++) acc[k] = 0.0;
  const int halves = mirror >= 0 ? 2 : 1;
  for (int e = body_start[bm] + t; e < body_start[bm + 1]; e += 256) {
    const double ds = m.area[e];
    for (int ip = 0; ip < halves; ip++) {
      const double r1 = (ip == 1 && mirror == 0) ? -1.0 : 1.0, r2 = (ip == 1 && mirror == 1) ? -1.0 : 1.0;
      const double sg[6] = {r1, r2, 1.0, r2, r1, r1 * r2};
      double2 xv = (ip == 0 ? X : Xm)[(size_t)r * n + e];
      cd phi{xv.x, xv.y};
      if (!rad && w.isol == 1) phi = phi + d_vinp(r1 * m.cx[e], r2 * m.cy[e], m.cz[e], f, w, w.beta[r - nd]);
      for (int k = 0; k < 6; k++) {
        double nk = sg[k] * m.nrm[(size_t)k * n + e];
        if (rad) {
          acc[k] += w.rho * phi.r * ds * nk;            // Im(i rho phi DS n_k)
          acc[6 + k] += w.rho * f.w1 * phi.i * ds * nk;  // -Re(i rho w phi DS n_k)
        } else {
          acc[k] += phi.r * ds * nk;
          acc[6 + k] += phi.i * ds * nk;
        }
      }
    }
  }
  for (int k = 0; k < 12; k++) red[k][t] = acc[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s)
      for (int k = 0; k < 12; k++) red[k][t] += red[k][t + s];
    __syncthreads();
  }
  if (t < 6) {
    const int row = 6 * bm + t;
    if (rad) {
      out[(size_t)row + (size_t)nd * r] = red[t][0];
      out[(size_t)nd * nd + row + (size_t)nd * r] = red[6 + t][0];
    } else {
      // i w rho (a + i b) = (-w rho b, w rho a)
      const int beta = r - nd;
      const double wr = f.w1 * w.rho;
      double* ex = out + 2 * (size_t)nd * nd + 2 * ((size_t)row + (size_t)nd * beta);
      ex[0] = -(wr * red[6 + t][0]);
      ex[1] = wr * red[t][0];
    }
  }
}

// ---------------------------------------------------------------------------
// Field points (CalPotentialMulti, PressureElevationMulti.f90:55-275, NSYS = 1).
// The assembly kernels above fill, for P field points, M = [S_fp | T_fp]
// (P x 2N): the single-layer entries DUM(1) and the double-layer entries
// DUM(2) of every (field point, hull panel) pair.  With Z = [W; -X] (2N x
// nrhs) -- W(j, 6b+m) = n_m(j) on body b's panels (FACTOR_NORMAL, :420-440),
// W(j, 6nb+k) = -DINCP(j, k), X the solved potentials MXPOT -- the sum over
// JEL of TERM1 - TERM2 (:151-200) is the product M Z.
// ---------------------------------------------------------------------------
// With a symmetry plane the mirror half's column set takes the mode signs of
// the plane (flip: bit m set when mode m + 1 changes sign, the SX / SY factors
// of :133-185), DINCP at the mirrored sources and the mirror half's MXPOT.
__global__ void k_field_weights(MeshDev m, int nbody, const double2* __restrict__ dincp,
                                const double2* __restrict__ X, double2* __restrict__ Z, int flip) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, col = blockIdx.y, n = m.n, nd = 6 * nbody;
  if (j >= n) return;
  double2 w;
  if (col < nd) {
    const double sg = (flip >> (col % 6)) & 1 ? -1.0 : 1.0;
    w = make_double2(m.body[j] == col / 6 ? sg * m.nrm[(size_t)(col % 6) * n + j] : 0.0, 0.0);
  } else {
    const double2 d = dincp[(size_t)(col - nd) * n + j];
    w = make_double2(-d.x, -d.y);
  }
  const size_t o = (size_t)col * 2 * n;
  const double2 x = X[(size_t)col * n + j];
  Z[o + j] = w;
  Z[o + n + j] = make_double2(-x.x, -x.y);
}

// The free-term test (:200-249): onb[p] = 1 when field point p lies on a hull
// panel -- within MAX(1E-6, 1E-3 PNSZ) of its plane, measured from its first
// vertex, and inside its polygon: the edge products (edge x (x - vertex)) . n
// all >= -TOLA or all <= TOLA, TOLA = 1E-3 PNSZ^2.  One thread per (panel,
// point); vtx[(3q + d) n + j] is coordinate d of vertex q of panel j; onb
// starts at zero.
__global__ __launch_bounds__(256) void k_on_body(MeshDev m, const double* __restrict__ vtx, int nfp,
                                                 const double* __restrict__ px, const double* __restrict__ py,
                                                 const double* __restrict__ pz, int* __restrict__ onb) {
#pragma clang fp contract(off)
  const int j = blockIdx.x * 256 + threadIdx.x, p = blockIdx.y;
  if (j >= m.n || p >= nfp) return;
  const int n = m.n, nc = m.ncn[j];
  const double x0 = px[p], x1 = py[p], x2 = pz[p];
  const double e1 = m.nrm[j], e2 = m.nrm[n + j], e3 = m.nrm[2 * n + j], pn = m.pnsz[j];
  const double tolp = fmax(1.0e-6, 1.0e-3 * pn), tola = 1.0e-3 * (pn * pn);
  const double dpln = (x0 - vtx[j]) * e1 + (x1 - vtx[(size_t)n + j]) * e2 + (x2 - vtx[2 * (size_t)n + j]) * e3;
  if (fabs(dpln) > tolp) return;
  double smin = 1.0e30, smax = -1.0e30;
  for (int q = 0; q < nc; q++) {
    const int r = q + 1 < nc ? q + 1 : 0;
    const double a0 = vtx[(size_t)(3 * q) * n + j], a1 = vtx[(size_t)(3 * q + 1) * n + j],
                 a2 = vtx[(size_t)(3 * q + 2) * n + j];
    const double ev0 = vtx[(size_t)(3 * r) * n + j] - a0, ev1 = vtx[(size_t)(3 * r + 1) * n + j] - a1,
                 ev2 = vtx[(size_t)(3 * r + 2) * n + j] - a2;
    const double wv0 = x0 - a0, wv1 = x1 - a1, wv2 = x2 - a2;
    const double cx = ev1 * wv2 - ev2 * wv1, cy = ev2 * wv0 - ev0 * wv2, cz = ev0 * wv1 - ev1 * wv0;
    const double sv = cx * e1 + cy * e2 + cz * e3;
    smin = fmin(smin, sv);
    smax = fmax(smax, sv);
  }
  if (smin >= -tola || smax <= tola) atomicOr(&onb[p], 1);
}

__global__ void k_fill_int(int* __restrict__ a, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) a[i] = 0;
}

// The factor of CalPressureMulti (i w rho) or CalElevationMulti (i w / g),
// rho and 1 at the zero- and infinite-frequency limits (:303-348).
__device__ __forceinline__ cd field_factor(cd v, int kind, double z, bool limit, double w1, double rho) {
  if (kind == 3) kind = fabs(z) > HX_F(1.0e-6) ? 1 : 2;  // the output drivers' choice (:446, :495)
  if (kind == 1) return limit ? cd{rho * v.r, rho * v.i} : cd{-((w1 * rho) * v.i), (w1 * rho) * v.r};
  if (kind == 2) return limit ? v : cd{-((w1 / kG) * v.i), (w1 / kG) * v.r};
  return v;
}

// POT = (M Z)(p, col) / SLD, SLD = 2 pi on a panel and 4 pi in the fluid
// (:251-258), plus VINP for a diffraction column when ISOL = 1 (:260-271),
// times the factor of `kind`; out[col * ld + p] for the chunk's points, and
// inc[k * ld + p] (may be null) the incident-wave value of heading k
// (CalPotentialIncMulti with its pressure/elevation factors, :277-300,
// :516-556).
__global__ void k_field_finish(int nfp, const double* __restrict__ px, const double* __restrict__ py,
                               const double* __restrict__ pz, const int* __restrict__ onb, FreqDev f, WaveDev w,
                               int nd, int kind, const double2* __restrict__ P, size_t ld, double2* __restrict__ out,
                               double2* __restrict__ inc) {
  f = freq_now(f);
  const int p = blockIdx.x * blockDim.x + threadIdx.x, col = blockIdx.y;
  if (p >= nfp) return;
  const bool limit = fabs(f.v) < HX_F(1.0e-8) || fabs(f.v + 1.0) < HX_F(1.0e-8);
  const double sld = onb[p] ? 2.0 * kPi : 4.0 * kPi;
  const double2 pv = P[(size_t)col * nfp + p];
  cd v{pv.x / sld, pv.y / sld};
  cd f0 = mk(0.0);
  if (col >= nd && w.isol == 1 && !limit) f0 = d_vinp(px[p], py[p], pz[p], f, w, w.beta[col - nd]);
  v = field_factor(v + f0, kind, pz[p], limit, f.w1, w.rho);
  out[(size_t)col * ld + p] = make_double2(v.r, v.i);
  if (inc && col >= nd) {
    const cd fi = field_factor(f0, kind, pz[p], limit, f.w1, w.rho);
    inc[(size_t)(col - nd) * ld + p] = make_double2(fi.r, fi.i);
  }
}

// ---------------------------------------------------------------------------
// Motions (SolveMotionMulti, SolveMotionMulti.f90:46-150): one workgroup per
// heading solves
//   [-w^2 (M + A) + i w (B + B_lin + V) + C + K] x = -Im(F) - i Re(F)
// by partial-pivoting elimination (ZGESV: the pivot is the first maximum of
// |re| + |im|), with V = 0 first and then, when the quadratic damping is not
// negligible, V(i,j) = B_quad(i,j) w |x_j| until sum_j |dx_j| / |x_j| <= 1e-6.
// coef: the force reductions (amas, bdmp: nd x nd; exfc: nd x nbeta complex);
// bm: the block-diagonal body matrices MATX, BLNR, BQDR, CRS, KSTF (nd x nd
// each); work: nd x nd complex per heading.  iters[b] = the steps taken, -1
// when kMotionMaxIt did not converge.
// ---------------------------------------------------------------------------
constexpr int kMotionMaxDof = 192, kMotionMaxIt = 1000;

__device__ __forceinline__ double cabs_d(double2 v) { return hypot(v.x, v.y); }
__device__ __forceinline__ double2 cdiv_d(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ double2 cmul_d(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ __launch_bounds__(256) void k_motion(int nd, double w1, const double* __restrict__ coef,
                                                const double* __restrict__ bm, int quad, double2* __restrict__ work,
                                                double2* __restrict__ dspl, int* __restrict__ iters) {
  __shared__ double2 x[kMotionMaxDof], r[kMotionMaxDof];
  __shared__ double red[256];
  __shared__ int redi[256];
  const int b = blockIdx.x, t = threadIdx.x;
  const size_t nn = (size_t)nd * nd;
  double2* L = work + (size_t)b * nn;
  const double *amas = coef, *bdmp = coef + nn;
  const double2* exfc = reinterpret_cast<const double2*>(coef + 2 * nn);
  const double *mass = bm, *blnr = bm + nn, *bqdr = bm + 2 * nn, *crs = bm + 3 * nn, *kstf = bm + 4 * nn;
  int steps = 0;
  for (int pass = 0;; pass++) {
    for (size_t o = t; o < nn; o += 256) {
      const int j = (int)(o / nd);
      double damp = bdmp[o] + blnr[o];
      if (pass > 0) damp = damp + bqdr[o] * w1 * cabs_d(x[j]);
      const double re = -(w1 * w1 * (mass[o] + amas[o])) + crs[o] + kstf[o];
      L[o] = make_double2(re, w1 * damp);
    }
    if (t < nd) {
      const double2 e = exfc[(size_t)b * nd + t];
      r[t] = make_double2(-e.y, -e.x);
    }
    __syncthreads();
    // factorisation with row interchanges applied to r as well
    for (int k = 0; k < nd; k++) {
      double v = -1.0;
      int idx = nd;
      if (k + t < nd) {
        const double2 c = L[(size_t)k * nd + k + t];
        v = fabs(c.x) + fabs(c.y);
        idx = k + t;
      }
      red[t] = v;
      redi[t] = idx;
      __syncthreads();
      for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (t < s2) {
          const double a = red[t], c = red[t + s2];
          if (c > a || (c == a && redi[t + s2] < redi[t])) {
            red[t] = c;
            redi[t] = redi[t + s2];
          }
        }
        __syncthreads();
      }
      const int p = redi[0];
      const bool zero = red[0] == 0.0;
      __syncthreads();
      if (p != k) {
        for (int c = t; c < nd; c += 256) {
          const double2 u = L[(size_t)c * nd + k];
          L[(size_t)c * nd + k] = L[(size_t)c * nd + p];
          L[(size_t)c * nd + p] = u;
        }
        if (t == 0) {
          const double2 u = r[k];
          r[k] = r[p];
          r[p] = u;
        }
      }
      __syncthreads();
      if (zero) continue;  // singular column: left as it is (ZGETRF reports INFO and goes on)
      const double2 rp = cdiv_d(make_double2(1.0, 0.0), L[(size_t)k * nd + k]);
      for (int q = k + 1 + t; q < nd; q += 256) L[(size_t)k * nd + q] = cmul_d(L[(size_t)k * nd + q], rp);
      __syncthreads();
      const int m = nd - k - 1;
      for (int o = t; o < m * m; o += 256) {
        const int q = k + 1 + o % m, c = k + 1 + o / m;
        const double2 l = L[(size_t)k * nd + q], u = L[(size_t)c * nd + k], a = L[(size_t)c * nd + q];
        const double2 lu = cmul_d(l, u);
        L[(size_t)c * nd + q] = make_double2(a.x - lu.x, a.y - lu.y);
      }
      __syncthreads();
    }
    for (int k = 0; k < nd; k++) {  // L y = P r (unit lower)
      if (t > k && t < nd) {
        const double2 lu = cmul_d(L[(size_t)k * nd + t], r[k]);
        r[t] = make_double2(r[t].x - lu.x, r[t].y - lu.y);
      }
      __syncthreads();
    }
    for (int k = nd - 1; k >= 0; k--) {  // U x = y
      if (t == 0) r[k] = cdiv_d(r[k], L[(size_t)k * nd + k]);
      __syncthreads();
      if (t < k) {
        const double2 lu = cmul_d(L[(size_t)k * nd + t], r[k]);
        r[t] = make_double2(r[t].x - lu.x, r[t].y - lu.y);
      }
      __syncthreads();
    }
    if (pass == 0) {
      if (t < nd) x[t] = r[t];
      __syncthreads();
      if (!quad) break;
      continue;
    }
    if (t == 0) {
      double rerr = 0.0;
      for (int j = 0; j < nd; j++) {
        const double a = cabs_d(r[j]);
        if (a > 0.0) rerr += cabs_d(make_double2(r[j].x - x[j].x, r[j].y - x[j].y)) / a;
      }
      red[0] = rerr;
    }
    __syncthreads();
    const double rerr = red[0];
    if (t < nd) x[t] = r[t];
    steps = pass;
    __syncthreads();
    if (rerr <= 1.0e-6) break;
    if (pass >= kMotionMaxIt) {
      steps = -1;
      break;
    }
  }
  if (t < nd) dspl[(size_t)b * nd + t] = x[t];
  if (t == 0) iters[b] = steps;
}

// ---------------------------------------------------------------------------
// Geometric symmetry (NSYS = 2): with T2 / S2 the entries of the mirrored
// field points (CalGreenFuncMulti.f90:141-177, IS = 2), the two systems of
// ASSB_LEFT_MULTI (AssbMatxMulti.f90:66-80; RXY = [1 1; 1 -1]):
// A1 = A + T2, A2 = A - T2 (A holds the 2 pi diagonal), and the single-layer
// combinations S+ = S + S2, S- = S - S2 that carry the right-hand sides.
// ---------------------------------------------------------------------------
__global__ void k_sym_combine(double2* __restrict__ A, double2* __restrict__ T2, double2* __restrict__ S,
                              double2* __restrict__ S2, size_t count) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += (size_t)gridDim.x * blockDim.x) {
    const double2 a = A[k], t = T2[k], s = S[k], u = S2[k];
    A[k] = make_double2(a.x + t.x, a.y + t.y);
    T2[k] = make_double2(a.x - t.x, a.y - t.y);
    S[k] = make_double2(s.x + u.x, s.y + u.y);
    S2[k] = make_double2(s.x - u.x, s.y - u.y);
  }
}

// DINCP at the mirror image of panel j: DINP at the mirrored centroid, the
// mirrored normal component negated (DBC_MULTI_RIGHT, BodyIntgrMulti.f90:150-170
// with IS != IP).  mirror 0: x -> -x, 1: y -> -y.
__global__ void k_dum2_mirror(MeshDev m, FreqDev f, WaveDev w, int mirror, double2* dum2) {
  f = freq_now(f);
  const int j = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (j >= m.n) return;
  const double sx = mirror == 0 ? -1.0 : 1.0, sy = mirror == 1 ? -1.0 : 1.0;
  cd d[3];
  d_dinp(sx * m.cx[j], sy * m.cy[j], m.cz[j], f, w, w.beta[b], d);
  const cd v = (sx * d[0]) * m.nrm[j] + (sy * d[1]) * m.nrm[m.n + j] + d[2] * m.nrm[2 * m.n + j];
  dum2[(size_t)b * m.n + j] = make_double2(v.r, v.i);
}

// The right-hand-side generators of the two systems (ASSB_RBC/DBC_MULTI with
// the RXY combination, AssbMatxMulti.f90:124-251): radiation mode m of body b
// is symmetric (system 1) or antisymmetric (system 2) under the mirror --
// x-mirror: sway, heave, roll symmetric; y-mirror: surge, heave, pitch -- and
// enters its system with weight 2 n_m on body b's panels; heading k enters
// system 1 with -(DINCP + DINCP_mirror) and system 2 with -(DINCP - DINCP_mirror).
__global__ void k_sym_weights(MeshDev m, int nbody, int nbeta, int mirror, const double2* __restrict__ dinc,
                              const double2* __restrict__ dincm, double2* __restrict__ W1,
                              double2* __restrict__ W2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, col = blockIdx.y, n = m.n, nd = 6 * nbody;
  if (j >= n) return;
  double2 w1, w2;
  if (col < nd) {
    const int md = col % 6;
    const bool sym = mirror == 0 ? (md == 1 || md == 2 || md == 3) : (md == 0 || md == 2 || md == 4);
    const double v = m.body[j] == col / 6 ? 2.0 * m.nrm[(size_t)md * n + j] : 0.0;
    w1 = make_double2(sym ? v : 0.0, 0.0);
    w2 = make_double2(sym ? 0.0 : v, 0.0);
  } else {
    const double2 a = dinc[(size_t)(col - nd) * n + j], b = dincm[(size_t)(col - nd) * n + j];
    w1 = make_double2(-(a.x + b.x), -(a.y + b.y));
    w2 = make_double2(-(a.x - b.x), -(a.y - b.y));
  }
  W1[(size_t)col * n + j] = w1;
  W2[(size_t)col * n + j] = w2;
}

// MXPOT(:, col, IP) = (X1 +- X2) / 2 (RADIATION/DIFFRACTION_SOLVER_MULTI,
// AssbMatxMulti.f90:306-316, :373-381): X <- the mesh half, X2 <- the mirror.
__global__ void k_sym_recombine(double2* __restrict__ X, double2* __restrict__ X2, size_t count) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += (size_t)gridDim.x * blockDim.x) {
    const double2 a = X[k], b = X2[k];
    X[k] = make_double2((a.x + b.x) / 2.0, (a.y + b.y) / 2.0);
    X2[k] = make_double2((a.x - b.x) / 2.0, (a.y - b.y) / 2.0);
  }
}

// The mirror images of count points (mirror 0: x -> -x, 1: y -> -y).
__global__ void k_reflect_points(int count, int mirror, const double* __restrict__ px, const double* __restrict__ py,
                                 const double* __restrict__ pz, double* __restrict__ qx, double* __restrict__ qy,
                                 double* __restrict__ qz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  qx[k] = mirror == 0 ? -px[k] : px[k];
  qy[k] = mirror == 1 ? -py[k] : py[k];
  qz[k] = pz[k];
}

__global__ void k_clear_flag(int* flags) { flags[0] = 0; }

namespace launch {

void assemble(hipStream_t s, const MeshDev& m, const FreqDev& f, double2* A, double2* S, int* flags) {
  dim3 grid((m.nf - m.row0 + 63) / 64, (m.n + 3) / 4);
  HX_LAUNCH(k_clear_flag, dim3(1), dim3(1), 0, s, flags);
  if (f.h < 0.0)
    HX_LAUNCH(k_assemble<true>, grid, dim3(256), 0, s, m, f, A, S, flags);
  else
    HX_LAUNCH(k_assemble<false>, grid, dim3(256), 0, s, m, f, A, S, flags);
}

void mode_tables(hipStream_t s, const MeshDev& m, const FreqDev& f) {
  HX_LAUNCH(k_mode_tables, dim3((m.nf + 255) / 256, kModes), dim3(256), 0, s, m, f);
}

void bin_count(hipStream_t s, const MeshDev& m, double h, int* count) {
  dim3 grid((m.nf - m.row0 + 63) / 64, (m.n + 3) / 4);
  HX_LAUNCH(k_bin_count, grid, dim3(256), 0, s, m, h, count);
}

void bin_scatter(hipStream_t s, const MeshDev& m, double h, int* cursor, unsigned* list) {
  dim3 grid((m.nf - m.row0 + 63) / 64, (m.n + 3) / 4);
  HX_LAUNCH(k_bin_scatter, grid, dim3(256), 0, s, m, h, cursor, list);
}

void clear_flags(hipStream_t s, int* flags) { HX_LAUNCH(k_clear_flag, dim3(1), dim3(1), 0, s, flags); }

void assemble_region(hipStream_t s, const MeshDev& m, const FreqDev& f, int region, bool paired,
                     const unsigned* list, size_t count, double2* A, double2* S, int* flags) {
  if (count == 0) return;
  const dim3 grid((unsigned)((count + 255) / 256));
  if (paired) {
    switch (region) {
      case kRegPidcock:
        HX_LAUNCH(k_assemble_pairs<kRegPidcock>, grid, dim3(256), 0, s, m, f, list, (int)count, A, S);
        break;
      case kRegEigene:
        HX_LAUNCH(k_assemble_pairs<kRegEigene>, grid, dim3(256), 0, s, m, f, list, (int)count, A, S);
        break;
      case kRegEigen:
        HX_LAUNCH(k_assemble_pairs<kRegEigen>, grid, dim3(256), 0, s, m, f, list, (int)count, A, S);
        break;
      default:
        break;  // LINTON pairs are always single entries
    }
    return;
  }
  switch (region) {
    case kRegLinton:
      HX_LAUNCH(k_assemble_list<kRegLinton>, grid, dim3(256), 0, s, m, f, list, (int)count, A, S, flags);
      break;
    case kRegPidcock:
      HX_LAUNCH(k_assemble_list<kRegPidcock>, grid, dim3(256), 0, s, m, f, list, (int)count, A, S, flags);
      break;
    case kRegEigene:
      HX_LAUNCH(k_assemble_list<kRegEigene>, grid, dim3(256), 0, s, m, f, list, (int)count, A, S, flags);
      break;
    default:
      HX_LAUNCH(k_assemble_list<kRegEigen>, grid, dim3(256), 0, s, m, f, list, (int)count, A, S, flags);
      break;
  }
}

void rhs(hipStream_t s, const MeshDev& m, const FreqDev& f, const WaveDev& w, int nbody, const int* body_start_h,
         const int* body_start_d, const double2* S, double2* dum2, double2* partial, int nchunk,
         const int* chunk_body_d, const int* chunk_j0_d, const int* chunk_j1_d, double2* X) {
  (void)body_start_h;
  (void)body_start_d;
  const int n = m.n;
  if (w.nbeta > 0) HX_LAUNCH(k_dum2, dim3((n + 255) / 256, w.nbeta), dim3(256), 0, s, m, f, w, dum2);
  int ngroups = w.nbeta > 0 ? (w.nbeta + kBetaGroup - 1) / kBetaGroup : 1;
  HX_LAUNCH(k_rhs_partial, dim3((m.nf + 63) / 64, nchunk, ngroups), dim3(64), 0, s, n, m.nf, m.nrm, S, dum2,
            w.nbeta, chunk_j0_d, chunk_j1_d, partial);
  HX_LAUNCH(k_rhs_reduce, dim3((m.nf + 255) / 256, 6 * nbody + w.nbeta), dim3(256), 0, s, m.nf, nbody, w.nbeta,
            nchunk, chunk_body_d, partial, X);
}

void dincp(hipStream_t s, const MeshDev& m, const FreqDev& f, const WaveDev& w, double2* out, int mirror) {
  if (w.nbeta == 0) return;
  if (mirror < 0)
    HX_LAUNCH(k_dum2, dim3((m.n + 255) / 256, w.nbeta), dim3(256), 0, s, m, f, w, out);
  else
    HX_LAUNCH(k_dum2_mirror, dim3((m.n + 255) / 256, w.nbeta), dim3(256), 0, s, m, f, w, mirror, out);
}

void reflect_points(hipStream_t s, int count, int mirror, const double* px, const double* py, const double* pz,
                    double* qx, double* qy, double* qz) {
  if (count > 0)
    HX_LAUNCH(k_reflect_points, dim3((count + 255) / 256), dim3(256), 0, s, count, mirror, px, py, pz, qx, qy, qz);
}

void field_weights(hipStream_t s, const MeshDev& m, int nbody, int nbeta, const double2* dincp, const double2* X,
                   double2* Z, int flip) {
  HX_LAUNCH(k_field_weights, dim3((m.n + 255) / 256, 6 * nbody + nbeta), dim3(256), 0, s, m, nbody, dincp, X, Z,
            flip);
}

void on_body(hipStream_t s, const MeshDev& m, const double* vtx, int nfp, const double* px, const double* py,
             const double* pz, int* onb, bool accumulate) {
  if (!accumulate) HX_LAUNCH(k_fill_int, dim3((nfp + 255) / 256), dim3(256), 0, s, onb, nfp);
  HX_LAUNCH(k_on_body, dim3((m.n + 255) / 256, nfp), dim3(256), 0, s, m, vtx, nfp, px, py, pz, onb);
}

void field_finish(hipStream_t s, int nfp, const double* px, const double* py, const double* pz, const int* onb,
                  const FreqDev& f, const WaveDev& w, int nd, int kind, const double2* P, size_t ld, double2* out,
                  double2* inc) {
  HX_LAUNCH(k_field_finish, dim3((nfp + 255) / 256, nd + w.nbeta), dim3(256), 0, s, nfp, px, py, pz, onb, f, w, nd,
            kind, P, ld, out, inc);
}

void motions(hipStream_t s, int nd, int nbeta, double w1, const double* coef, const double* bm, int quad,
             double2* work, double2* dspl, int* iters) {
  if (nbeta > 0) HX_LAUNCH(k_motion, dim3(nbeta), dim3(256), 0, s, nd, w1, coef, bm, quad, work, dspl, iters);
}

void forces(hipStream_t s, const MeshDev& m, const FreqDev& f, const WaveDev& w, int nbody, const int* body_start_d,
            const double2* X, double* out, const double2* Xm, int mirror) {
  HX_LAUNCH(k_forces, dim3(nbody, 6 * nbody + w.nbeta), dim3(256), 0, s, m, f, w, nbody, body_start_d, X, Xm,
            mirror, out);
}

void sym_combine(hipStream_t s, double2* A, double2* T2, double2* S, double2* S2, size_t count) {
  HX_LAUNCH(k_sym_combine, dim3(4096), dim3(256), 0, s, A, T2, S, S2, count);
}

void sym_weights(hipStream_t s, const MeshDev& m, const FreqDev& f, const WaveDev& w, int nbody, int mirror,
                 double2* dinc, double2* W1, double2* W2) {
  if (w.nbeta > 0) {
    HX_LAUNCH(k_dum2, dim3((m.n + 255) / 256, w.nbeta), dim3(256), 0, s, m, f, w, dinc);
    HX_LAUNCH(k_dum2_mirror, dim3((m.n + 255) / 256, w.nbeta), dim3(256), 0, s, m, f, w, mirror,
              dinc + (size_t)m.n * w.nbeta);
  }
  HX_LAUNCH(k_sym_weights, dim3((m.n + 255) / 256, 6 * nbody + w.nbeta), dim3(256), 0, s, m, nbody, w.nbeta, mirror,
            dinc, dinc + (size_t)m.n * w.nbeta, W1, W2);
}

void sym_recombine(hipStream_t s, double2* X, double2* X2, size_t count) {
  HX_LAUNCH(k_sym_recombine, dim3(1024), dim3(256), 0, s, X, X2, count);
}

}  // namespace launch
}  // namespace hamsx
