// sg.hpp -- self-gravity of the polar method on a logarithmic grid
// (compute_selfgravity, src/SgMain.cu:22-47, with ComputeFFTKernel,
// ComputeConvolutionKernel, ComputeSgAccKernel and Update_sgvelocityKernel,
// src/Kernels.cu:949-1041).
//
// The reference runs six 2-D C2C transforms of the zero-padded
// (2*NRAD) x NSEC grid per substep.  Here the two real inputs
// Sr = Sigma*sqrt(r/r0) and St = Sr*r/r0 travel as one complex field
// z = Sr + i*St, and the two real outputs g_r, g_t come back as one complex
// field, so a substep needs one forward and one inverse transform, in three
// LDS-resident passes (FP64 throughout, whatever the field precision):
//   A  row transforms of the NRAD non-zero rows of z;
//   B  per pair of columns (k, NSEC-k): column transforms, the Hermitian
//      split Sr^ = (Z(k) + conj Z(-k))/2, St^ = (Z(k) - conj Z(-k))/(2i),
//      W^ = -G (Kr^ Sr^ + i Kt^ St^) with the kernel transforms precomputed
//      once per disk on the host, inverse column transforms (rows < NRAD);
//   C  inverse row transforms, normalisation, the self-cell term, the ring
//      mean of g_r (sub-Keplerian boundary) and the fields SG_Accr, SG_Acct.
// Transforms are mixed-radix (4, 2, 3, 5) Stockham, the factorisation and
// twiddle tables being per disk (host: csrc/sg.cpp).
#pragma once
#include "kernels.hpp"

namespace fgx {

__device__ __forceinline__ cxd cx_add(cxd a, cxd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cxd cx_sub(cxd a, cxd b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cxd cx_mul(cxd a, cxd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cxd cx_scale(cxd a, double s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ cxd cx_i(cxd a, double s) { return {-s * a.y, s * a.x}; }  // (i*s)*a

template <int R>
__device__ __forceinline__ void dft_small(cxd (&v)[R], double s) {
  if constexpr (R == 2) {
    const cxd a = cx_add(v[0], v[1]), b = cx_sub(v[0], v[1]);
    v[0] = a; v[1] = b;
  } else if constexpr (R == 4) {
    const cxd a0 = cx_add(v[0], v[2]), a1 = cx_sub(v[0], v[2]);
    const cxd b0 = cx_add(v[1], v[3]), b1 = cx_i(cx_sub(v[1], v[3]), s);
    v[0] = cx_add(a0, b0); v[2] = cx_sub(a0, b0); v[1] = cx_add(a1, b1); v[3] = cx_sub(a1, b1);
  } else if constexpr (R == 3) {
    const double sg = s * 0.86602540378443864676;
    const cxd t1 = cx_add(v[1], v[2]), d = cx_i(cx_sub(v[1], v[2]), sg);
    const cxd m = cx_add(v[0], cx_scale(t1, -0.5));
    v[0] = cx_add(v[0], t1); v[1] = cx_add(m, d); v[2] = cx_sub(m, d);
  } else {
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    const double s1 = s * 0.95105651629515357212, s2 = s * 0.58778525229247312917;
    const cxd t1 = cx_add(v[1], v[4]), t2 = cx_add(v[2], v[3]), t3 = cx_sub(v[1], v[4]), t4 = cx_sub(v[2], v[3]);
    const cxd m1 = cx_add(v[0], cx_add(cx_scale(t1, c1), cx_scale(t2, c2)));
    const cxd m2 = cx_add(v[0], cx_add(cx_scale(t1, c2), cx_scale(t2, c1)));
    const cxd d1 = cx_i(cx_add(cx_scale(t3, s1), cx_scale(t4, s2)), 1.0);
    const cxd d2 = cx_i(cx_sub(cx_scale(t3, s2), cx_scale(t4, s1)), 1.0);
    v[0] = cx_add(v[0], cx_add(t1, t2)); v[1] = cx_add(m1, d1); v[4] = cx_sub(m1, d1);
    v[2] = cx_add(m2, d2); v[3] = cx_sub(m2, d2);
  }
}

// LDS position of element e of a transform buffer: one 16-byte pad after
// every eight elements, so that the stores of the first stages (lanes R
// elements apart) fall in different banks; a buffer of N elements takes
// lds_len(N) slots
__device__ __host__ __forceinline__ int lp(int e) { return e + (e >> 3); }
__device__ __host__ __forceinline__ int lds_len(int n) { return n + (n >> 3) + 1; }

// one radix-R Stockham stage over nl lines of length n stored back to back
// (element positions lp())
template <int R>
__device__ __forceinline__ void fft_stage(const cxd* src, cxd* dst, int n, int nl, int Ns, double s,
                                          const cxd* __restrict__ tw) {
  const int m = n / R;
  const int ts = n / (Ns * R);
  for (int t = threadIdx.x; t < m * nl; t += blockDim.x) {
    const int l = t / m, j = t - l * m;
    const int e0 = l * n;
    const int k = j % Ns;
    cxd v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v[r] = src[lp(e0 + j + r * m)];
      if (r > 0) {
        cxd w = tw[r * k * ts];   // exp(-2 pi i r k / (Ns R))
        if (s > 0) w.y = -w.y;
        v[r] = cx_mul(v[r], w);
      }
    }
    dft_small<R>(v, s);
    const int base = e0 + (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[lp(base + r * Ns)] = v[r];
  }
}

// nl transforms of length n (factors fac[0..nf)), sign s = -1 forward, +1
// inverse (unnormalised); a holds the input, b is scratch; returns the buffer
// holding the result.  Ends with a barrier.
__device__ inline cxd* lds_fft(cxd* a, cxd* b, int n, int nl, const int* fac, int nf, double s,
                               const cxd* __restrict__ tw) {
  cxd* src = a;
  cxd* dst = b;
  int Ns = 1;
  for (int st = 0; st < nf; ++st) {
    const int R = fac[st];
    if (R == 4) fft_stage<4>(src, dst, n, nl, Ns, s, tw);
    else if (R == 2) fft_stage<2>(src, dst, n, nl, Ns, s, tw);
    else if (R == 3) fft_stage<3>(src, dst, n, nl, Ns, s, tw);
    else fft_stage<5>(src, dst, n, nl, Ns, s, tw);
    __syncthreads();
    cxd* x = src; src = dst; dst = x;
    Ns *= R;
  }
  return src;
}

// Block sizes: the passes are launched with 256 to 1024 threads, as many as a
// stage has butterflies (a transform's LDS footprint holds a block per CU at
// the largest grids, so its waves are all the CU has); every loop over
// elements strides by the block size, and no value depends on it.
// pass A: one block per ring i < NRAD
template <class real>
__global__ void __launch_bounds__(1024) k_sg_rows_fwd(Dev<real> D) {
  extern __shared__ unsigned char fgx_smem[];
  FGX_RING_PROLOGUE
  if (i >= nr) return;
  cxd* a = reinterpret_cast<cxd*>(fgx_smem);
  cxd* b = a + lds_len(ns);
  const double q = RT(rmed_d)[rb + i] / RT(rmed_d)[rb];
  const double sq = sqrt(q);
  for (int j = threadIdx.x; j < ns; j += blockDim.x) {
    const double var = s_whole_d(D, rb, i, (real)AT(D.dens, i, j)) * sq;
    a[lp(j)] = cxd{var, var * q};
  }
  __syncthreads();
  const cxd* out = lds_fft(a, b, ns, 1, c.fac_ns, c.nf_ns, -1.0, D.sgtw + c.tw_ns);
  cxd* w = D.sgw + c.sgw_o + (long long)i * ns;
  for (int j = threadIdx.x; j < ns; j += blockDim.x) w[j] = out[lp(j)];
}

// pass B: one block per pair of columns (k, NSEC-k) of a disk
template <class real>
__global__ void __launch_bounds__(1024) k_sg_cols(Dev<real> D) {
  extern __shared__ unsigned char fgx_smem[];
  const RingTile t = D.sg_ctiles[xcd_tile((int)blockIdx.x, (int)gridDim.x)];
  if (t.d < 0) return;
  const DiskConst& c = D.dc[t.d];
  const DiskState& s = D.st[t.d];
  if (skip_disk(D, s, t.d)) return;
  const int ns = c.nsec, nr = c.nrad, n = 2 * nr;
  const int ka = t.i, kb = (ns - ka) % ns;
  const int nl = (ka == kb) ? 1 : 2;
  cxd* a = reinterpret_cast<cxd*>(fgx_smem);
  cxd* b = a + lds_len(2 * n);
  const cxd* w = D.sgw + c.sgw_o;
  for (int u = threadIdx.x; u < nl * n; u += blockDim.x) {
    const int l = u / n, ki = u - l * n;
    a[lp(u)] = (ki < nr) ? w[(long long)ki * ns + (l ? kb : ka)] : cxd{0.0, 0.0};
  }
  __syncthreads();
  const cxd* tw = D.sgtw + c.tw_nr;
  cxd* z = lds_fft(a, b, n, nl, c.fac_nr, c.nf_nr, -1.0, tw);
  cxd* o = (z == a) ? b : a;
  const cxd* kr = D.sgkr + c.sgk_o;
  const cxd* kt = D.sgkt + c.sgk_o;
  for (int u = threadIdx.x; u < nl * n; u += blockDim.x) {
    const int l = u / n, ki = u - l * n;
    const int col = l ? kb : ka;
    const int pl = (nl == 1) ? 0 : 1 - l;
    const int kim = (ki == 0) ? 0 : n - ki;
    const cxd zz = z[lp(u)], zp = z[lp(pl * n + kim)];
    const cxd zc = cxd{zp.x, -zp.y};
    const cxd sr = cx_scale(cx_add(zz, zc), 0.5);
    const cxd d = cx_sub(zz, zc);
    const cxd st = cxd{0.5 * d.y, -0.5 * d.x};   // d / (2i)
    const long long kk = (long long)col * n + ki;   // the kernel transforms are stored by column
    const cxd gr = cx_mul(kr[kk], sr), gt = cx_mul(kt[kk], st);
    o[lp(u)] = cx_scale(cx_add(gr, cx_i(gt, 1.0)), -kG);
  }
  __syncthreads();
  const cxd* res = lds_fft(o, z, n, nl, c.fac_nr, c.nf_nr, 1.0, tw);
  cxd* wo = D.sgw + c.sgw_o;
  for (int u = threadIdx.x; u < nl * n; u += blockDim.x) {
    const int l = u / n, ki = u - l * n;
    if (ki < nr) wo[(long long)ki * ns + (l ? kb : ka)] = res[lp(u)];
  }
}

// pass C: one block per ring i < NRAD
template <class real>
__global__ void __launch_bounds__(1024) k_sg_rows_inv(Dev<real> D) {
  extern __shared__ unsigned char fgx_smem[];
  __shared__ double redd[1024 / 64];
  FGX_RING_PROLOGUE
  if (i >= nr) return;
  cxd* a = reinterpret_cast<cxd*>(fgx_smem);
  cxd* b = a + lds_len(ns);
  const cxd* w = D.sgw + c.sgw_o + (long long)i * ns;
  for (int j = threadIdx.x; j < ns; j += blockDim.x) a[lp(j)] = w[j];
  __syncthreads();
  const cxd* out = lds_fft(a, b, ns, 1, c.fac_ns, c.nf_ns, 1.0, D.sgtw + c.tw_ns);
  const double q = RT(rmed_d)[rb + i] / RT(rmed_d)[rb];
  const double norm = c.sg_rstep * c.sg_tstep / ((double)(2 * nr) * (double)ns);
  const double nrr = norm / sqrt(q), ntt = norm / (q * sqrt(q));
  const double self = kG * c.sg_rstep * c.sg_tstep / c.sg_eps;
  // the fields and the ring mean of g_r are formed by the first FGX_TPB
  // threads, sector j by thread j mod FGX_TPB, and summed over their four
  // waves in order, whatever the block size, so no bit depends on it
  double acc = 0;
  if (threadIdx.x < FGX_TPB) {
    for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
      const cxd o = out[lp(j)];
      const double gr = o.x * nrr + self * s_whole_d(D, rb, i, (real)AT(D.dens, i, j));
      AT(D.sg_accr, i, j) = (real)gr;
      AT(D.sg_acct, i, j) = (real)(o.y * ntt);
      acc += gr;
    }
  }
  acc = wave_sum(acc);
  __syncthreads();
  if ((threadIdx.x & 63) == 0 && threadIdx.x < FGX_TPB) redd[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0;
    for (int k = 0; k < FGX_TPB / 64; ++k) r += redd[k];
    D.axisg[rb + i] = r / (double)ns;   // Make1Dprofile(2)
  }
}

// Update_sgvelocityKernel (Kernels.cu:1024-1041) on VradInt / VthetaInt
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_sg_update(Dev<real> D) {
  FGX_STRIP_BEGIN {
    const double dt = s.dt;
    if (i > 0)
      AT(D.vradint, i, j) = (real)sg_vrad_new((double)AT(D.vradint, i, j), dt, RT(rinf_d)[rb + i], RT(rmed_d)[rb + i - 1],
                                              RT(rmed_d)[rb + i], (double)AT(D.sg_accr, i, j),
                                              (double)AT(D.sg_accr, i - 1, j), RT(invdiffrmed_d)[rb + i]);
    // (j-1)%nsec at j = 0 is -1 in the reference: linear index i*nsec-1, the
    // previous ring's last sector (periodic on ring 0, where it would read
    // before the array; ext_fix_sg_wrap makes every ring periodic)
    double am;
    if (j > 0) am = (double)AT(D.sg_acct, i, j - 1);
    else if (D.fix_sgwrap || i == 0) am = (double)AT(D.sg_acct, i, ns - 1);
    else am = (double)AT(D.sg_acct, i - 1, ns - 1);
    AT(D.vthetaint, i, j) = (real)sg_vtheta_new((double)AT(D.vthetaint, i, j), dt, (double)AT(D.sg_acct, i, j), am);
  }
}

}  // namespace fgx
