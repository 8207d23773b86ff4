// sg.cpp -- host side of the self-gravity solver (polar method on a
// logarithmic grid, src/SgMain.cu): the FP64 transforms of the two Green's
// functions, done once per disk, and the initial acceleration field that
// set-up needs (Init_planetarysys_withSG, Init_azimutalvelocity_withSG).
//
// The FFT is a mixed-radix (4, 2, 3, 5) Stockham autosort transform written
// from the standard decimation-in-time recurrence: stage s with radix R and
// Ns = product of the earlier radices maps x[j + r*N/R] (r < R), twiddled by
// exp(sign*2*pi*i*r*(j mod Ns)/(Ns*R)), through an R-point DFT to
// y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns].  The device kernels use the same
// factorisation and the same twiddle table, exp(-2*pi*i*m/n) (conjugated for
// the inverse).  The 2-D transforms and the Green's functions of a large grid
// (2048 x 3072 points at NRAD=1024) run over the host's cores: lines and rows
// are independent, so the result does not depend on the thread count.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

#include "fgx_internal.h"

namespace fgx {

using cd = std::complex<double>;

int sg_factor(int n, int* fac, int* nf) {
  *nf = 0;
  if (n < 1) return FGX_EPARAM;
  const int pref[] = {4, 2, 3, 5};
  for (int r : pref)
    while (n % r == 0 && *nf < FGX_SG_MAXFAC) { fac[(*nf)++] = r; n /= r; }
  return n == 1 ? FGX_OK : FGX_EPARAM;
}

namespace {

void dft_small(cd* v, int R, int sign) {
  const double s = sign;
  if (R == 2) {
    const cd a = v[0] + v[1], b = v[0] - v[1];
    v[0] = a; v[1] = b;
  } else if (R == 4) {
    const cd is(0, s);
    const cd a0 = v[0] + v[2], a1 = v[0] - v[2], b0 = v[1] + v[3], b1 = (v[1] - v[3]) * is;
    v[0] = a0 + b0; v[2] = a0 - b0; v[1] = a1 + b1; v[3] = a1 - b1;
  } else if (R == 3) {
    const double c = -0.5, sg = s * std::sqrt(3.0) / 2.0;
    const cd t1 = v[1] + v[2], d = (v[1] - v[2]) * cd(0, sg);
    const cd m = v[0] + c * t1;
    v[0] = v[0] + t1; v[1] = m + d; v[2] = m - d;
  } else {  // 5
    const double c1 = std::cos(2 * M_PI / 5), c2 = std::cos(4 * M_PI / 5);
    const double s1 = s * std::sin(2 * M_PI / 5), s2 = s * std::sin(4 * M_PI / 5);
    const cd t1 = v[1] + v[4], t2 = v[2] + v[3], t3 = v[1] - v[4], t4 = v[2] - v[3];
    const cd m1 = v[0] + c1 * t1 + c2 * t2, m2 = v[0] + c2 * t1 + c1 * t2;
    const cd d1 = cd(0, 1) * (s1 * t3 + s2 * t4), d2 = cd(0, 1) * (s2 * t3 - s1 * t4);
    v[0] = v[0] + t1 + t2; v[1] = m1 + d1; v[4] = m1 - d1; v[2] = m2 + d2; v[3] = m2 - d2;
  }
}

// f(lo, hi) over [0, n) in contiguous chunks on up to 16 threads
template <class F>
void parallel_for(int n, F f) {
  const int nt = std::max(1, std::min({16, (int)std::thread::hardware_concurrency(), n / 16}));
  if (nt <= 1) { f(0, n); return; }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) th.emplace_back(f, (int)((long long)n * t / nt), (int)((long long)n * (t + 1) / nt));
  for (std::thread& x : th) x.join();
}

// twiddle table of a length-n transform: w[m] = exp(-2*pi*i*m/n)
std::vector<cd> twiddles(int n) {
  std::vector<cd> w(n);
  for (int m = 0; m < n; ++m) {
    const double a = 2.0 * M_PI * (double)m / (double)n;
    w[m] = cd(std::cos(a), -std::sin(a));
  }
  return w;
}

// one line of length n (contiguous in a), unnormalised, sign -1 forward;
// tw = twiddles(n)
void fft_line(cd* a, cd* b, int n, const int* fac, int nf, int sign, const cd* tw) {
  cd* src = a;
  cd* dst = b;
  int Ns = 1;
  cd v[5];
  for (int st = 0; st < nf; ++st) {
    const int R = fac[st];
    const int m = n / R;
    const int ts = n / (Ns * R);
    for (int j = 0; j < m; ++j) {
      const int k = j % Ns;
      v[0] = src[j];
      for (int r = 1; r < R; ++r) {
        const cd w = tw[r * k * ts];   // exp(-2 pi i r k / (Ns R))
        v[r] = src[j + r * m] * (sign > 0 ? std::conj(w) : w);
      }
      dft_small(v, R, sign);
      const int base = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
    }
    std::swap(src, dst);
    Ns *= R;
  }
  if (src != a) std::memcpy(a, src, sizeof(cd) * n);
}

// in-place 2-D transform of a rows x cols array (row-major); only rows
// [0, rows_in) of the input are non-zero (the row pass skips the others,
// whose transform is zero) and only rows [0, rows_out) of the output are kept
// (the inverse row pass skips the others)
void fft2d(std::vector<cd>& x, int rows, int cols, int sign, int rows_in, int rows_out) {
  int fr[FGX_SG_MAXFAC], fc[FGX_SG_MAXFAC], nr = 0, nc = 0;
  sg_factor(rows, fr, &nr);
  sg_factor(cols, fc, &nc);
  const std::vector<cd> twr = twiddles(rows), twc = twiddles(cols);
  // forward: rows, then columns; inverse: columns, then rows (the kept ones)
  auto row_pass = [&](int nrows) {
    parallel_for(nrows, [&](int lo, int hi) {
      std::vector<cd> tmp(cols);
      for (int i = lo; i < hi; ++i) fft_line(&x[(size_t)i * cols], tmp.data(), cols, fc, nc, sign, twc.data());
    });
  };
  // columns eight at a time, so that each row's cache line is read and
  // written once per group rather than once per column
  constexpr int G = 8;
  auto col_pass = [&]() {
    parallel_for((cols + G - 1) / G, [&](int lo, int hi) {
      std::vector<cd> line((size_t)G * rows), tmp(rows);
      for (int g = lo; g < hi; ++g) {
        const int j0 = g * G, nj = std::min(G, cols - j0);
        for (int i = 0; i < rows; ++i)
          for (int q = 0; q < nj; ++q) line[(size_t)q * rows + i] = x[(size_t)i * cols + j0 + q];
        for (int q = 0; q < nj; ++q) fft_line(&line[(size_t)q * rows], tmp.data(), rows, fr, nr, sign, twr.data());
        for (int i = 0; i < rows; ++i)
          for (int q = 0; q < nj; ++q) x[(size_t)i * cols + j0 + q] = line[(size_t)q * rows + i];
      }
    });
  };
  if (sign < 0) {
    row_pass(rows_in);
    col_pass();
  } else {
    col_pass();
    row_pass(rows_out);
  }
}
void fft2d(std::vector<cd>& x, int rows, int cols, int sign) { fft2d(x, rows, cols, sign, rows, rows); }

}  // namespace

int sg_host_setup(HostDisk& h, std::vector<double>& khat_r, std::vector<double>& khat_t,
                  std::vector<double>& axisg) {
  const fgx_params& P = h.P;
  const int NR = h.nrad, NS = h.nsec, NR2 = 2 * NR;
  int f[FGX_SG_MAXFAC], nf;
  if (sg_factor(NS, f, &nf) || sg_factor(NR2, f, &nf))
    return set_error(FGX_EPARAM, "self-gravity needs NSEC and 2*NRAD with prime factors 2, 3 and 5 only "
                                 "(NSEC=%d, NRAD=%d)", NS, NR);
  // SGP_eps, SGP_rstep, SGP_tstep (src/Main.cu:308-310); the softening key
  // defaults to the planets' THICKNESSSMOOTHING as in the reference
  const double soft = P.ext_sg_softening > 0 ? P.ext_sg_softening : P.thicknesssmoothing;
  h.sg_eps = soft * (double)P.aspectratio;
  h.sg_rstep = std::log(h.radii[NR] / h.radii[0]) / (double)NR;
  h.sg_tstep = 2.0 * M_PI / (double)NS;
  const double eps = h.sg_eps;
  // Green's functions on the log grid (compute_kernel, SgMain.cu:49-81), FP64
  std::vector<cd> kr((size_t)NR2 * NS), kt((size_t)NR2 * NS);
  // the angle's cosine and sine per sector, the radial terms per row
  std::vector<double> cth(NS), sth(NS);
  for (int j = 0; j < NS; ++j) {
    const double th = 2.0 * M_PI * (double)j / (double)NS;
    cth[j] = std::cos(th);
    sth[j] = std::sin(th);
  }
  parallel_for(NR2, [&](int lo, int hi) {
    for (int i = lo; i < hi; ++i) {
      const double u = (i < NR) ? std::log(h.radii[i] / h.radii[0]) : -std::log(h.radii[NR2 - i] / h.radii[0]);
      const double eu = eps * eps * std::exp(u), ch = std::cosh(u), emu = std::exp(-u);
      for (int j = 0; j < NS; ++j) {
        const double x = eu + 2.0 * (ch - cth[j]);
        const double base = 1.0 / (x * std::sqrt(x));   // x^-1.5
        kr[(size_t)i * NS + j] = (1.0 + eps * eps - cth[j] * emu) * base;
        kt[(size_t)i * NS + j] = sth[j] * base;
      }
    }
  });
  fft2d(kr, NR2, NS, -1);
  fft2d(kt, NR2, NS, -1);
  // stored by column (entry (i, j) at j * 2NRAD + i): the device reads them a
  // column pair at a time
  khat_r.resize(2 * kr.size());
  khat_t.resize(2 * kt.size());
  parallel_for((NS + 7) / 8, [&](int lo, int hi) {   // eight columns at a time
    for (int g = lo; g < hi; ++g) {
      const int j0 = 8 * g, nj = std::min(8, NS - j0);
      for (int i = 0; i < NR2; ++i)
        for (int jj = 0; jj < nj; ++jj) {
          const size_t k = (size_t)i * NS + j0 + jj, q = (size_t)(j0 + jj) * NR2 + i;
          khat_r[2 * q] = kr[k].real(); khat_r[2 * q + 1] = kr[k].imag();
          khat_t[2 * q] = kt[k].real(); khat_t[2 * q + 1] = kt[k].imag();
        }
    }
  });
  // radial acceleration of the initial density (ComputeFFTKernel,
  // ComputeConvolutionKernel, ComputeSgAccKernel; Kernels.cu:949-1021)
  std::vector<cd> sr((size_t)NR2 * NS, 0.0);
  parallel_for(NR, [&](int lo, int hi) {
    for (int i = lo; i < hi; ++i)
      for (int j = 0; j < NS; ++j) sr[(size_t)i * NS + j] = h.dens[(size_t)i * NS + j] * std::sqrt(h.rmed[i] / h.rmed[0]);
  });
  fft2d(sr, NR2, NS, -1, NR, NR2);   // rows >= NRAD of the input are zero
  parallel_for(NR2, [&](int lo, int hi) {
    for (size_t k = (size_t)lo * NS; k < (size_t)hi * NS; ++k) sr[k] = -kG * kr[k] * sr[k];
  });
  fft2d(sr, NR2, NS, +1, NR2, NR);   // rows < NRAD of the result are read
  const double norm = h.sg_rstep * h.sg_tstep / ((double)NR2 * (double)NS);
  axisg.assign(NR + 1, 0.0);
  for (int i = 0; i < NR; ++i) {
    const double nrm = norm / std::sqrt(h.rmed[i] / h.rmed[0]);
    double s = 0;
    for (int j = 0; j < NS; ++j)
      s += sr[(size_t)i * NS + j].real() * nrm + kG * h.dens[(size_t)i * NS + j] * h.sg_rstep * h.sg_tstep / eps;
    axisg[i] = s / NS;  // Make1Dprofile(2)
  }
  // Init_planetarysys_withSG (SgMain.cu:166-195): planets on circular orbits
  // that feel the disk's radial self-gravity at their semi-major axis
  for (int k = 0; k < h.sys.nb; ++k) {
    const double r = h.sys.x[k];
    const double dist = (double)(float)(r / (1. + P.eccentricity));
    int ipl = 0;
    while (ipl < NR - 2 && h.rmed[ipl] <= dist) ipl++;
    const double ri = h.rmed[ipl], rip1 = h.rmed[ipl + 1], dr = rip1 - ri;
    const double sgacc = ((dist - ri) * axisg[ipl + 1] + (rip1 - dist) * axisg[ipl]) / dr;
    h.sys.vy[k] *= std::sqrt(1. - dist * dist * sgacc);
  }
  return FGX_OK;
}

}  // namespace fgx
