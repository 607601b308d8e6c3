// noise_fft.hip — FFT-filtered correlated noise on MI355X (gfx950).
//
// Implements the circulant-embedding filter of Percival (2006) the way the
// reference applies it (noise/hops_noise.py:630-716):
//   tilde c_j = c_j exp(-2 pi i nu j)          (rotate so c[n_t-1] is real)
//   s_w = Re FFT([tilde c, conj(reverse(tilde c[1:-1]))]), negatives -> 0
//   eta = unrotate( IFFT( FFT(z) * sqrt(s_w / 2) )[:n_t] )
// One workgroup owns one row and keeps it in LDS (rows longer than the LDS
// holds work in a library-owned global scratch row instead, which the same
// workgroup barriers order).  The transform length is
// N = 2(n_t - 1).  Powers of two run a radix-2 decimation-in-frequency
// forward pass (natural -> bit-reversed order) and a decimation-in-time
// inverse pass (bit-reversed -> natural), so the spectral multiply happens in
// bit-reversed order and no permutation pass is needed.  Other lengths use
// Bluestein's chirp-z algorithm on a power-of-two length M >= 2N-1 with the
// same two passes.  All arithmetic is FP64; the output is complex64 like the
// reference's noise.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>

#include "../../include/hops_mi355x.h"

extern "C" int hops_internal_set_error(const char *msg);

namespace {

constexpr int kMaxLds = 8192;     // complex doubles held in LDS (128 KiB)
constexpr int kMaxLen = 1 << 22;  // longest transform (global scratch rows)
constexpr int kThreads = 256;

__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 twiddle(double sgn, long j, long len) {
  double s, c;
  sincospi(sgn * 2.0 * (double)j / (double)len, &s, &c);
  return make_double2(c, s);
}
__device__ __forceinline__ unsigned bitrev(unsigned x, int logm) {
  return __brev(x) >> (32 - logm);
}

// In-place radix-2 DIF: natural order in, bit-reversed order out.
__device__ void fft_dif(double2 *a, int M, double sgn) {
  for (int len = M; len >= 2; len >>= 1) {
    const int half = len >> 1;
    for (int t = threadIdx.x; t < (M >> 1); t += blockDim.x) {
      const int blk = t / half, j = t - blk * half;
      const int i0 = blk * len + j, i1 = i0 + half;
      const double2 u = a[i0], v = a[i1];
      a[i0] = make_double2(u.x + v.x, u.y + v.y);
      a[i1] = cmul2(make_double2(u.x - v.x, u.y - v.y), twiddle(sgn, j, len));
    }
    __syncthreads();
  }
}

// In-place radix-2 DIT: bit-reversed order in, natural order out.
__device__ void fft_dit(double2 *a, int M, double sgn) {
  for (int len = 2; len <= M; len <<= 1) {
    const int half = len >> 1;
    for (int t = threadIdx.x; t < (M >> 1); t += blockDim.x) {
      const int blk = t / half, j = t - blk * half;
      const int i0 = blk * len + j, i1 = i0 + half;
      const double2 u = a[i0];
      const double2 v = cmul2(a[i1], twiddle(sgn, j, len));
      a[i0] = make_double2(u.x + v.x, u.y + v.y);
      a[i1] = make_double2(u.x - v.x, u.y - v.y);
    }
    __syncthreads();
  }
}

// Bluestein chirp w_j = exp(i pi j^2 / N)  (j^2 reduced mod 2N).
__device__ __forceinline__ double2 chirp(long j, long N) {
  const long q = (j * j) % (2 * N);
  double s, c;
  sincospi((double)q / (double)N, &s, &c);
  return make_double2(c, s);
}

// Forward DFT of length N held in a[0..N) (natural order); M = 2^logm.
// pow2 (M == N): result left in BIT-REVERSED order.  Bluestein (M > N): result
// in natural order in a[0..N).  Bhat_d: FFT_M of the chirp filter in
// bit-reversed order (Bluestein only).
__device__ void dft_forward(double2 *a, int N, int M, int logm, const double2 *Bhat_d) {
  if (M == N) {
    fft_dif(a, M, -1.0);
    return;
  }
  for (int j = threadIdx.x; j < M; j += blockDim.x) {
    if (j < N) {
      const double2 w = chirp(j, N);
      a[j] = cmul2(a[j], make_double2(w.x, -w.y));
    } else {
      a[j] = make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  fft_dif(a, M, -1.0);
  for (int j = threadIdx.x; j < M; j += blockDim.x) a[j] = cmul2(a[j], Bhat_d[j]);
  __syncthreads();
  fft_dit(a, M, 1.0);
  const double inv = 1.0 / (double)M;
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    const double2 w = chirp(k, N);
    const double2 v = cmul2(a[k], make_double2(w.x, -w.y));
    a[k] = make_double2(v.x * inv, v.y * inv);
  }
  __syncthreads();
}

// FFT_M of the Bluestein filter b_j = w_j (j < N), b_{M-j} = w_j (0 < j < N),
// stored in bit-reversed order.
__global__ __launch_bounds__(kThreads) void k_bluestein_filter(double2 *Bhat, int N, int M, int in_place) {
  extern __shared__ double2 lds[];
  double2 *a = in_place ? Bhat : lds;
  for (int j = threadIdx.x; j < M; j += blockDim.x) {
    double2 v = make_double2(0.0, 0.0);
    if (j < N) v = chirp(j, N);
    else if (j > M - N) v = chirp(M - j, N);
    a[j] = v;
  }
  __syncthreads();
  fft_dif(a, M, -1.0);
  if (!in_place)
    for (int j = threadIdx.x; j < M; j += blockDim.x) Bhat[j] = lds[j];
}

// Embedding spectrum of each correlation function row.
__global__ __launch_bounds__(kThreads) void k_embedding(const float *__restrict__ ang, const float *__restrict__ mag,
                                                        int n_t, int N, int M, int logm, const double2 *Bhat,
                                                        double *s_w, double *nu, double2 *gscratch) {
  extern __shared__ double2 lds_[];
  const int row = blockIdx.x;
  double2 *lds = gscratch ? gscratch + (long)row * M : lds_;
  const float *ar = ang + (long)row * n_t;
  const float *mr = mag + (long)row * n_t;
  // rotation so that the last point is real
  const double nur = (double)ar[n_t - 1] / (((double)n_t - 1.0) * 2.0 * M_PI);
  for (int j = threadIdx.x; j < n_t; j += blockDim.x) {
    const double a = (double)ar[j], m = (double)mr[j];
    double s1, c1, s2, c2;
    sincos(a, &s1, &c1);
    sincospi(-2.0 * nur * (double)j, &s2, &c2);
    const double2 t = cmul2(make_double2(m * c1, m * s1), make_double2(c2, s2));
    lds[j] = t;
    if (j >= 1 && j <= n_t - 2) lds[N - j] = make_double2(t.x, -t.y);
  }
  __syncthreads();
  dft_forward(lds, N, M, logm, Bhat);
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    const int src = (M == N) ? (int)bitrev((unsigned)k, logm) : k;
    const double v = lds[src].x;
    s_w[(long)row * N + k] = v < 0.0 ? 0.0 : v;
  }
  if (threadIdx.x == 0) nu[row] = nur;
}

__device__ __forceinline__ void put_noise(float2 *out, long i, double2 v) {
  out[i] = make_float2((float)v.x, (float)v.y);
}
__device__ __forceinline__ void put_noise(double2 *out, long i, double2 v) { out[i] = v; }

// Correlated noise of each raw row; OutT float2 (the complex64 noise of the
// reference) or double2 (the complex128 values its cubic interpolation is
// built on).
template <typename OutT>
__global__ __launch_bounds__(kThreads) void k_filter(const float2 *__restrict__ z, int n_t, int N, int M,
                                                     int logm, const double2 *Bhat,
                                                     const double *__restrict__ s_w,
                                                     const double *__restrict__ nu,
                                                     const int32_t *__restrict__ row_lop,
                                                     OutT *__restrict__ out, double2 *gscratch) {
  extern __shared__ double2 lds_[];
  const int row = blockIdx.x;
  double2 *lds = gscratch ? gscratch + (long)row * M : lds_;
  const int l = row_lop[row];
  const float2 *zr = z + (long)row * N;
  const double *sw = s_w + (long)l * N;
  for (int j = threadIdx.x; j < N; j += blockDim.x) lds[j] = make_double2((double)zr[j].x, (double)zr[j].y);
  __syncthreads();
  dft_forward(lds, N, M, logm, Bhat);
  if (M == N) {
    // spectrum in bit-reversed order: multiply, then DIT inverse -> natural
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      const double f = sqrt(sw[bitrev((unsigned)i, logm)] / 2.0);
      lds[i] = make_double2(lds[i].x * f, lds[i].y * f);
    }
    __syncthreads();
    fft_dit(lds, M, 1.0);
  } else {
    // inverse DFT = conj(DFT(conj(X))) / N
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      const double f = sqrt(sw[k] / 2.0);
      lds[k] = make_double2(lds[k].x * f, -lds[k].y * f);
    }
    __syncthreads();
    dft_forward(lds, N, M, logm, Bhat);
    for (int k = threadIdx.x; k < N; k += blockDim.x) lds[k].y = -lds[k].y;
    __syncthreads();
  }
  const double inv = 1.0 / (double)N;
  const double nur = nu[l];
  for (int j = threadIdx.x; j < n_t; j += blockDim.x) {
    double s, c;
    sincospi(2.0 * nur * (double)j, &s, &c);  // conj(E2)
    const double2 y = make_double2(lds[j].x * inv, lds[j].y * inv);
    const double2 v = cmul2(y, make_double2(c, s));
    put_noise(out, (long)row * n_t + j, v);
  }
}

// Cubic interpolation of each complex128 row y[0..n) on the uniform grid
// x_i = i h, with not-a-knot ends (the spline scipy.interpolate.interp1d
// builds for kind="cubic"), evaluated at t_j = j dt for j < n_out.  With
// uniform spacing the not-a-knot rows fix M_1 = r_1 / 6 and M_{n-2} =
// r_{n-2} / 6 (r_i = 6 (y_{i+1} - 2 y_i + y_{i-1}) / h^2); the second
// derivatives M_2..M_{n-3} solve M_{i-1} + 4 M_i + M_{i+1} = r_i by the Thomas
// algorithm (cp: its shared pivots 1 / (4 - cp_{k-1})), and M_0 = 2 M_1 - M_2,
// M_{n-1} = 2 M_{n-2} - M_{n-3}.  Evaluated at t = t0 + j dt.  One thread per
// row; work: [R][n] complex128.
__global__ __launch_bounds__(kThreads) void k_spline(const double2 *__restrict__ y, int R, int n, double h,
                                                     double t0, double dt, int n_out,
                                                     const double *__restrict__ cp,
                                                     double2 *__restrict__ work, double2 *__restrict__ out) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= R) return;
  const double2 *yr = y + (long)row * n;
  double2 *M = work + (long)row * n;
  const double s6 = 6.0 / (h * h);
  auto rhs = [&](int i) {
    return make_double2(s6 * (yr[i + 1].x - 2.0 * yr[i].x + yr[i - 1].x),
                        s6 * (yr[i + 1].y - 2.0 * yr[i].y + yr[i - 1].y));
  };
  const double2 r1 = rhs(1), rl = rhs(n - 2);
  M[1] = make_double2(r1.x / 6.0, r1.y / 6.0);
  M[n - 2] = make_double2(rl.x / 6.0, rl.y / 6.0);
  const int m = n - 4;  // unknowns M_2 .. M_{n-3}
  if (m > 0) {
    // forward sweep: M[2 + k] holds d'_k
    for (int k = 0; k < m; ++k) {
      const int i = 2 + k;
      double2 r = rhs(i);
      if (k == 0) r = make_double2(r.x - M[1].x, r.y - M[1].y);
      if (k == m - 1) r = make_double2(r.x - M[n - 2].x, r.y - M[n - 2].y);
      const double2 prev = k == 0 ? make_double2(0.0, 0.0) : M[i - 1];
      const double piv = k == 0 ? 0.25 : cp[k];  // 1 / (4 - cp_{k-1})
      M[i] = make_double2((r.x - prev.x) * piv, (r.y - prev.y) * piv);
    }
    // back substitution: c'_k = cp[k] as the pivot of row k (c'_k = 1 * piv)
    for (int k = m - 2; k >= 0; --k) {
      const int i = 2 + k;
      const double c = k == 0 ? 0.25 : cp[k];
      M[i] = make_double2(M[i].x - c * M[i + 1].x, M[i].y - c * M[i + 1].y);
    }
  }
  M[0] = make_double2(2.0 * M[1].x - M[2].x, 2.0 * M[1].y - M[2].y);
  M[n - 1] = make_double2(2.0 * M[n - 2].x - M[n - 3].x, 2.0 * M[n - 2].y - M[n - 3].y);
  const double inv6h = 1.0 / (6.0 * h), h6 = h / 6.0;
  for (int j = 0; j < n_out; ++j) {
    const double t = t0 + (double)j * dt;
    int i = (int)floor(t / h);
    if (i > n - 2) i = n - 2;
    if (i < 0) i = 0;
    const double B = t - (double)i * h, A = h - B;
    const double a3 = A * A * A * inv6h, b3 = B * B * B * inv6h;
    const double2 y0 = yr[i], y1 = yr[i + 1], m0 = M[i], m1 = M[i + 1];
    out[(long)row * n_out + j] =
        make_double2(m0.x * a3 + m1.x * b3 + (y0.x / h - m0.x * h6) * A + (y1.x / h - m1.x * h6) * B,
                     m0.y * a3 + m1.y * b3 + (y0.y / h - m0.y * h6) * A + (y1.y / h - m1.y * h6) * B);
  }
}

// Sum of n float32 values a[0], a[stride], ... in NumPy's pairwise order for
// complex64 reductions (numpy/_core/src/umath/loops_utils.h.src): fewer than
// 4 values one after another starting from -0; 4..64 values in four running
// sums combined as (r0 + r1) + (r2 + r3), then the remainder one by one.
__device__ __forceinline__ float np_pairwise_sum(const float *a, int stride, int n) {
  if (n < 4) {
    float r = -0.0f;
    for (int i = 0; i < n; ++i) r += a[(long)i * stride];
    return r;
  }
  float r0 = a[0], r1 = a[stride], r2 = a[2L * stride], r3 = a[3L * stride];
  int i = 4;
  for (; i < n - (n % 4); i += 4) {
    r0 += a[(long)i * stride];
    r1 += a[(long)(i + 1) * stride];
    r2 += a[(long)(i + 2) * stride];
    r3 += a[(long)(i + 3) * stride];
  }
  float res = (r0 + r1) + (r2 + r3);
  for (; i < n; ++i) res += a[(long)i * stride];
  return res;
}

// out[row][i] = mean of z[row][i .. i+w) (complex64, NumPy's summation order
// and a float32 division by w); 0 where the window passes the end
__global__ __launch_bounds__(kThreads) void k_window_mean(const float2 *__restrict__ z, float2 *__restrict__ out,
                                                         int n_t, int w) {
#pragma clang fp contract(off)
  const int row = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_t) return;
  const float2 *zr = z + (long)row * n_t;
  float2 v = make_float2(0.f, 0.f);
  if (i + w <= n_t) {
    const float *base = reinterpret_cast<const float *>(zr + i);
    v.x = np_pairwise_sum(base, 2, w) / (float)w;
    v.y = np_pairwise_sum(base + 1, 2, w) / (float)w;
  }
  out[(long)row * n_t + i] = v;
}

struct Plan {
  int N, M, logm;
  double2 *Bhat;  // library-owned, one per length, freed at process exit
};

std::mutex g_plan_mu;
std::map<int, Plan> g_plans;

int set_lds(const void *fn, size_t bytes) {
  if (bytes > 64 * 1024 &&
      hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
    return hops_internal_set_error("hipFuncSetAttribute(MaxDynamicSharedMemorySize) failed");
  return 0;
}

// scratch rows for transforms longer than the LDS holds, one buffer per
// stream: transforms queued on different streams (host threads serving
// different batches) never share rows, and those on one stream are ordered
// by it.  Library-owned, grown on demand, kept until the process ends.
struct Scratch {
  double2 *p = nullptr;
  size_t bytes = 0;
};
std::mutex g_scratch_mu;
std::map<hipStream_t, Scratch> g_scratch;

int scratch_rows(size_t bytes, hipStream_t st, double2 **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  Scratch &s = g_scratch[st];
  if (bytes > s.bytes) {
    if (s.p) {
      // only this stream's work reads the old rows
      if (hipStreamSynchronize(st) != hipSuccess || hipFree(s.p) != hipSuccess)
        return hops_internal_set_error("releasing the noise FFT scratch failed");
      s.p = nullptr;
      s.bytes = 0;
    }
    if (hipMalloc(&s.p, bytes) != hipSuccess) {
      s.p = nullptr;
      return hops_internal_set_error("hipMalloc of the noise FFT scratch failed");
    }
    s.bytes = bytes;
  }
  *out = s.p;
  return 0;
}

int make_plan(int n_t, hipStream_t st, Plan *out) {
  if (n_t < 3) return hops_internal_set_error("noise time axis needs at least 3 points");
  const long N = 2L * (n_t - 1);
  long M = 1;
  int logm = 0;
  while (M < N) {
    M <<= 1;
    ++logm;
  }
  Plan p{(int)N, (int)M, logm, nullptr};
  if (M != N) {
    M = 1;
    logm = 0;
    while (M < 2 * N - 1) {
      M <<= 1;
      ++logm;
    }
    p.M = (int)M;
    p.logm = logm;
  }
  if (p.M > kMaxLen) return hops_internal_set_error("noise FFT length exceeds 2^22 points");
  if (p.M != p.N) {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(p.N);
    if (it != g_plans.end()) {
      *out = it->second;
      return 0;
    }
    if (hipMalloc(&p.Bhat, sizeof(double2) * p.M) != hipSuccess)
      return hops_internal_set_error("hipMalloc of the Bluestein filter failed");
    const bool big = p.M > kMaxLds;
    if (!big && set_lds((const void *)k_bluestein_filter, sizeof(double2) * p.M)) return -1;
    k_bluestein_filter<<<1, kThreads, big ? 0 : sizeof(double2) * p.M, st>>>(p.Bhat, p.N, p.M, big ? 1 : 0);
    if (hipGetLastError() != hipSuccess) return hops_internal_set_error("Bluestein filter launch failed");
    // a plan is shared by every stream: publish it once its filter is written
    if (hipStreamSynchronize(st) != hipSuccess) return hops_internal_set_error("Bluestein filter failed");
    g_plans[p.N] = p;
  }
  *out = p;
  return 0;
}


}  // namespace

extern "C" int hops_noise_embedding(const float *ang_d, const float *mag_d, int n_lop, int n_t, double *s_w_d,
                                    double *nu_d, void *stream) {
  if (n_lop <= 0) return 0;
  if (!ang_d || !mag_d || !s_w_d || !nu_d) return hops_internal_set_error("hops_noise_embedding: NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  Plan p;
  if (int rc = make_plan(n_t, st, &p)) return rc;
  double2 *gs = nullptr;
  if (p.M > kMaxLds) {
    if (int rc = scratch_rows(sizeof(double2) * p.M * (size_t)n_lop, st, &gs)) return rc;
  }
  const size_t lds = gs ? 0 : sizeof(double2) * p.M;
  if (int rc = set_lds((const void *)k_embedding, lds)) return rc;
  k_embedding<<<n_lop, kThreads, lds, st>>>(ang_d, mag_d, n_t, p.N, p.M, p.logm, p.Bhat, s_w_d, nu_d, gs);
  if (hipGetLastError() != hipSuccess) return hops_internal_set_error("k_embedding launch failed");
  return 0;
}

extern "C" int hops_noise_filter(const void *z_d, int R, int n_t, const double *s_w_d, const double *nu_d,
                                 const int32_t *row_lop_d, void *out_d, void *stream) {
  if (R <= 0) return 0;
  if (!z_d || !s_w_d || !nu_d || !row_lop_d || !out_d)
    return hops_internal_set_error("hops_noise_filter: NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  Plan p;
  if (int rc = make_plan(n_t, st, &p)) return rc;
  double2 *gs = nullptr;
  if (p.M > kMaxLds) {
    if (int rc = scratch_rows(sizeof(double2) * p.M * (size_t)R, st, &gs)) return rc;
  }
  const size_t lds = gs ? 0 : sizeof(double2) * p.M;
  if (int rc = set_lds((const void *)k_filter<float2>, lds)) return rc;
  k_filter<float2><<<R, kThreads, lds, st>>>((const float2 *)z_d, n_t, p.N, p.M, p.logm, p.Bhat, s_w_d, nu_d,
                                              row_lop_d, (float2 *)out_d, gs);
  if (hipGetLastError() != hipSuccess) return hops_internal_set_error("k_filter launch failed");
  return 0;
}

extern "C" int hops_noise_filter_c128(const void *z_d, int R, int n_t, const double *s_w_d, const double *nu_d,
                                      const int32_t *row_lop_d, void *out_d, void *stream) {
  if (R <= 0) return 0;
  if (!z_d || !s_w_d || !nu_d || !row_lop_d || !out_d)
    return hops_internal_set_error("hops_noise_filter_c128: NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  Plan p;
  if (int rc = make_plan(n_t, st, &p)) return rc;
  double2 *gs = nullptr;
  if (p.M > kMaxLds) {
    if (int rc = scratch_rows(sizeof(double2) * p.M * (size_t)R, st, &gs)) return rc;
  }
  const size_t lds = gs ? 0 : sizeof(double2) * p.M;
  if (int rc = set_lds((const void *)k_filter<double2>, lds)) return rc;
  k_filter<double2><<<R, kThreads, lds, st>>>((const float2 *)z_d, n_t, p.N, p.M, p.logm, p.Bhat, s_w_d, nu_d,
                                               row_lop_d, (double2 *)out_d, gs);
  if (hipGetLastError() != hipSuccess) return hops_internal_set_error("k_filter launch failed");
  return 0;
}

extern "C" int hops_noise_spline(const void *y_d, int R, int n, double h, double t0, double dt, int n_out,
                                 const double *cp_d, void *work_d, void *out_d, void *stream) {
  if (R <= 0 || n_out <= 0) return 0;
  if (!y_d || !cp_d || !work_d || !out_d) return hops_internal_set_error("hops_noise_spline: NULL pointer");
  if (n < 4) return hops_internal_set_error("cubic interpolation needs at least 4 noise points");
  if (!(h > 0.0) || !(dt > 0.0)) return hops_internal_set_error("hops_noise_spline: spacing must be positive");
  if (t0 < 0.0 || t0 + (double)(n_out - 1) * dt > (double)(n - 1) * h * (1.0 + 1e-12))
    return hops_internal_set_error("hops_noise_spline: evaluation outside the noise time axis");
  k_spline<<<(R + kThreads - 1) / kThreads, kThreads, 0, (hipStream_t)stream>>>(
      (const double2 *)y_d, R, n, h, t0, dt, n_out, cp_d, (double2 *)work_d, (double2 *)out_d);
  if (hipGetLastError() != hipSuccess) return hops_internal_set_error("k_spline launch failed");
  return 0;
}

extern "C" int hops_noise_window_mean(const void *z_d, void *out_d, int R, int n_t, int w, void *stream) {
  if (R <= 0) return 0;
  if (!z_d || !out_d) return hops_internal_set_error("hops_noise_window_mean: NULL pointer");
  if (w < 1 || w > 64) return hops_internal_set_error("noise averaging window must hold 1..64 samples");
  if (n_t < 1 || R > 65535) return hops_internal_set_error("hops_noise_window_mean: bad shape");
  dim3 grid((unsigned)((n_t + kThreads - 1) / kThreads), (unsigned)R);
  k_window_mean<<<grid, kThreads, 0, (hipStream_t)stream>>>((const float2 *)z_d, (float2 *)out_d, n_t, w);
  if (hipGetLastError() != hipSuccess) return hops_internal_set_error("k_window_mean launch failed");
  return 0;
}
