// hops_engine.hip — HOPS propagation on MI355X (gfx950).
//
// One trajectory's hierarchy vector Phi is a set of rows r = k*n_state + s
// (auxiliary k, state s).  A batch of B trajectories that share one basis
// keeps its pools trajectory-major, so the neighbour gathers of one
// trajectory stay inside one ~0.1-2 MB window; tiles of a trajectory are
// placed on one XCD (workgroup i runs on XCD i % 8) so that window lives in
// that XCD's L2.
//
// Per RK4 stage two launches run:
//   k_prologue  one workgroup per trajectory: <L>, z-memory stage value and
//               derivative, z-hat, the normalisation correction, the LTC
//               terms and the per-trajectory coefficients of the system block
//               (-iH + sum_j c_j L_j ...) — all in FP64.
//   k_stage     one thread per row: the derivative (aux-diagonal K0, system
//               block, ELL super-operator K2_kp1+K2_km1, Z2 term, -norm_corr)
//               fused with the RK4 combination.  Stages run in complex R
//               (double or float); the state carry Phi stays complex128.
// A LINEAR step computes its four prologues in one launch
// (k_prologue_linear4).  Bases of at most resident_rows rows per trajectory
// run a whole call in one launch instead (k_steps_resident); larger ones
// replay K captured steps from a HIP graph.
//
// Normalisation (NORMALIZED NONLINEAR) is lazy: the scale sigma = 1/|psi| of
// the state produced by the previous step is computed in FP64 by the stage-0
// prologue and folded into every read of Phi during the next step.
//
// Floating-point contraction is off for the whole file: a multiply and an add
// stay two roundings unless the code asks for fma() itself.  The launched
// stage forms, the resident form and the graph-replayed steps inline the same
// functions into different kernels; with contraction left to the compiler
// their FP64 prologue scalars could differ in the last bit between kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>
#include <algorithm>

#include "../../include/hops_mi355x.h"

#pragma clang fp contract(off)

namespace {

thread_local std::string g_err;

int fail(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail("%s failed: %s", #expr, hipGetErrorString(e_));            \
  } while (0)

template <typename R> struct alignas(2 * sizeof(R)) cx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ cx<R> mk(R a, R b) {
  cx<R> c;
  c.x = a;
  c.y = b;
  return c;
}
template <typename R>
__device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) {
  return mk<R>(a.x + b.x, a.y + b.y);
}
template <typename R>
__device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) {
  return mk<R>(a.x - b.x, a.y - b.y);
}
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a += b * c as four explicit fused multiply-adds: every kernel variant
// performs exactly this sequence, so they agree bit for bit.
__device__ __forceinline__ float fmaR(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fmaR(double a, double b, double c) { return fma(a, b, c); }
template <typename R>
__device__ __forceinline__ void cfma(cx<R> &a, cx<R> b, cx<R> c) {
  a.x = fmaR(b.x, c.x, a.x);
  a.x = fmaR(-b.y, c.y, a.x);
  a.y = fmaR(b.x, c.y, a.y);
  a.y = fmaR(b.y, c.x, a.y);
}
template <typename R>
__device__ __forceinline__ cx<R> cscale(cx<R> a, R s) {
  return mk<R>(a.x * s, a.y * s);
}
template <typename R> __device__ __forceinline__ cx<R> cconj(cx<R> a) {
  return mk<R>(a.x, -a.y);
}
// Loads and stores with the nontemporal hint, for the arrays a stage streams
// through once (RK4 accumulator, FP64 carry): they then displace less of the
// L2 that holds the neighbour rows the stage gathers.
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
  static_assert(sizeof(T) == 8 || sizeof(T) == 16, "8- or 16-byte elements");
  T v;
  if constexpr (sizeof(T) == 8) {
    const unsigned long long u = __builtin_nontemporal_load(reinterpret_cast<const unsigned long long *>(p));
    __builtin_memcpy(&v, &u, 8);
  } else {
    typedef double d2v __attribute__((ext_vector_type(2)));
    const d2v u = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(p));
    __builtin_memcpy(&v, &u, 16);
  }
  return v;
}
template <typename T>
__device__ __forceinline__ void st_stream(T *p, T v) {
  static_assert(sizeof(T) == 8 || sizeof(T) == 16, "8- or 16-byte elements");
  if constexpr (sizeof(T) == 8) {
    unsigned long long u;
    __builtin_memcpy(&u, &v, 8);
    __builtin_nontemporal_store(u, reinterpret_cast<unsigned long long *>(p));
  } else {
    typedef double d2v __attribute__((ext_vector_type(2)));
    d2v u;
    __builtin_memcpy(&u, &v, 16);
    __builtin_nontemporal_store(u, reinterpret_cast<d2v *>(p));
  }
}
template <int NT, typename T>
__device__ __forceinline__ T ld_s(const T *p) {
  if constexpr (NT != 0) return ld_stream(p);
  else return *p;
}
template <int NT, typename T>
__device__ __forceinline__ void st_s(T *p, T v) {
  if constexpr (NT != 0) st_stream(p, v);
  else *p = v;
}
// Complex division by Smith's (1962) scaled algorithm with a reciprocal scale,
// the same operation sequence NumPy uses for complex128 division, so that
// g/w reproduces the reference's super-operator values bit for bit.
template <typename R>
__device__ __forceinline__ cx<R> cdiv(cx<R> a, cx<R> b) {
#pragma clang fp contract(off)
  if (fabs(b.x) >= fabs(b.y)) {
    const R rat = b.y / b.x;
    const R scl = (R)1 / (b.x + b.y * rat);
    return mk<R>((a.x + a.y * rat) * scl, (a.y - a.x * rat) * scl);
  }
  const R rat = b.x / b.y;
  const R scl = (R)1 / (b.y + b.x * rat);
  return mk<R>((a.x * rat + a.y) * scl, (a.y * rat - a.x) * scl);
}
template <typename R, typename S>
__device__ __forceinline__ cx<R> ccast(cx<S> a) {
  return mk<R>((R)a.x, (R)a.y);
}
__device__ __forceinline__ cx<double> ld2(const double *p, int i) {
  return mk<double>(p[2 * i], p[2 * i + 1]);
}

typedef cx<double> zd;

// Block-wide sum of a complex double (blockDim.x == 256, 4 wavefronts of 64).
__device__ zd block_sum(zd v, double *red) {
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o, 64);
    v.y += __shfl_xor(v.y, o, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    red[2 * wid] = v.x;
    red[2 * wid + 1] = v.y;
  }
  __syncthreads();
  zd t = mk<double>(0.0, 0.0);
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) {
    t.x += red[2 * w];
    t.y += red[2 * w + 1];
  }
  __syncthreads();
  return t;
}

// Pools (phi, y0, ya, yb, acc) hold groups of T = 2^lane_shift trajectories
// interleaved row by row: element (b, r) lives at ((b >> lane_shift) * nr + r)
// * T + (b & (T - 1)).  A gathered cache line then carries T trajectories.
struct Lane {
  long base;
  int T;
};
__host__ __device__ __forceinline__ Lane lane_of(const hops_batch_t &bt, long nr, int b) {
  Lane l;
  l.T = 1 << bt.lane_shift;
  l.base = (long)(b >> bt.lane_shift) * nr * l.T + (b & (l.T - 1));
  return l;
}

// ---------------------------------------------------------------------------
// Super-operator build (eom/eom_hops_ksuper.py:36-230, :575-668).
// Row r = (k, s).  Up connection (k -> k+e_m):  Kp1[(k,s),(k+e_m,s')] =
// -(g_m/w_m) L_m[s,s'].  Down connection (k -> k-e_m, depth k_m of aux k):
// Km1[(k,s),(k-e_m,s_r)] = (k_m w_m) conj(L_m[s_r,s]).
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void build_ell_row(const hops_op_t &op, const long r) {
#pragma clang fp contract(off)
  const int n = op.n_state;
  const long nr = (long)n * op.n_hier;
  if (r >= nr) return;
  const int k = (int)(r / n), s = (int)(r - (long)k * n);
  cx<R> *val = reinterpret_cast<cx<R> *>(op.ell_val_d);
  int e = 0;
  for (int u = op.up_ptr_d[k]; u < op.up_ptr_d[k + 1]; ++u) {
    const int m = op.up_mode_d[u], k2 = op.up_aux_d[u];
    const int j = op.mode_lrel_d[m];
    const zd gw = cdiv(ld2(op.mode_g_d, m), ld2(op.mode_w_d, m));
    const zd ngw = mk<double>(-gw.x, -gw.y);
    const int q0 = op.lr_ptr_d[j * (n + 1) + s], q1 = op.lr_ptr_d[j * (n + 1) + s + 1];
    for (int q = q0; q < q1; ++q) {
      if (e >= op.ell_w) return;  // host sized ell_w; never taken
      const zd v = cmul(ngw, ld2(op.lr_val_d, q));
      op.ell_col_d[(long)e * nr + r] = k2 * n + op.lr_col_d[q];
      val[(long)e * nr + r] = ccast<R>(v);
      ++e;
    }
  }
  for (int u = op.dn_ptr_d[k]; u < op.dn_ptr_d[k + 1]; ++u) {
    const int m = op.dn_mode_d[u], k2 = op.dn_aux_d[u];
    const int j = op.mode_lrel_d[m];
    const zd w = ld2(op.mode_w_d, m);
    const double kv = (double)op.dn_val_d[u];
    const zd kw = mk<double>(kv * w.x, kv * w.y);
    const int q0 = op.lc_ptr_d[j * (n + 1) + s], q1 = op.lc_ptr_d[j * (n + 1) + s + 1];
    for (int q = q0; q < q1; ++q) {
      if (e >= op.ell_w) return;
      const zd v = cmul(kw, cconj(ld2(op.lc_val_d, q)));
      op.ell_col_d[(long)e * nr + r] = k2 * n + op.lc_row_d[q];
      val[(long)e * nr + r] = ccast<R>(v);
      ++e;
    }
  }
  for (; e < op.ell_w; ++e) {
    op.ell_col_d[(long)e * nr + r] = -1;
    val[(long)e * nr + r] = mk<R>((R)0, (R)0);
  }
}

// K0[k] = -sum_m k_m W_m over the absolute modes of aux k
// (eom/eom_hops_ksuper.py:36-90).
__device__ __forceinline__ void build_k0_aux(const hops_op_t &op, const int k) {
#pragma clang fp contract(off)
  if (k >= op.n_hier) return;
  double sx = 0.0, sy = 0.0;
  for (int q = op.aux_ptr_d[k]; q < op.aux_ptr_d[k + 1]; ++q) {
    const int m = op.aux_mode_d[q];
    const double v = (double)op.aux_val_d[q];
    sx += v * op.abs_w_d[2 * m];
    sy += v * op.abs_w_d[2 * m + 1];
  }
  op.k0_d[2 * k] = -sx;
  op.k0_d[2 * k + 1] = -sy;
}

template <typename R>
__global__ __launch_bounds__(256) void k_build_ell(const hops_op_t op) {
  build_ell_row<R>(op, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_build_k0(const hops_op_t op) {
  build_k0_aux(op, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Per-trajectory coefficient layout (doubles):
//   A  [n_pat] complex   system block for every auxiliary
//   A0 [n_pat] complex   system block for the physical wave function
//   zc [n_mode] complex  (g_m/w_m) conj(<L_m>) for the Z2 term
//   nc complex           normalisation correction
//   sigma                scale of Phi for this step
// Workspace (doubles): kz [n_mode_abs] complex, zacc [n_mode_abs] complex.
// ---------------------------------------------------------------------------
struct CoefLayout {
  int A, A0, zc, nc, sig;
};
__host__ __device__ inline CoefLayout coef_layout(const hops_op_t &op) {
  CoefLayout c;
  c.A = 0;
  c.A0 = 2 * op.n_pat;
  c.zc = 4 * op.n_pat;
  c.nc = 4 * op.n_pat + 2 * op.n_mode;
  c.sig = c.nc + 2;
  return c;
}

__constant__ double c_rk[4] = {0.0, 0.5, 0.5, 1.0};

// Stage prologue: eom/hops_eom.py:329-441 (noise terms, <L>, norm_corr, LTC),
// eom/eom_functions.py:11-309, integrator/integrator_rk.py:53-64 (z_mem stage
// values, noise index map [0,1,1,2]).
// One workgroup per trajectory b; stage < 0 evaluates the derivative at yin
// with the current z-memory and no RK bookkeeping (adaptive error estimate).
template <typename R>
__device__ __forceinline__ void prologue_body(const hops_op_t &op, const hops_batch_t &bt, const int b,
                                              int stage, int tidx, const cx<R> *__restrict__ yin,
                                              int store_slot) {
  const bool eval = stage < 0;
  extern __shared__ double sm[];
  const int n = op.n_state, na = op.n_act, Ma = op.n_mode_abs;
  const int tid = threadIdx.x, nt = blockDim.x;
  const long nr = (long)n * op.n_hier;
  const Lane ln = lane_of(bt, nr, b);
  const CoefLayout L = coef_layout(op);
  double *coef = bt.coef_d + (long)b * bt.coef_stride;
  double *ws = bt.ws_d + (long)b * bt.ws_stride;
  double *zmem = bt.zmem_d + (long)b * 2 * Ma;
  const bool nonlin = bt.eom != HOPS_EOM_LINEAR;
  const bool normed = bt.eom == HOPS_EOM_NORMALIZED_NONLINEAR;

  double *psi = sm;                 // 2n
  double *avg = psi + 2 * n;        // 2na
  double *avq = avg + 2 * na;       // 2na
  double *cj = avq + 2 * na;        // 2na  noise coefficient
  double *zh = cj + 2 * na;         // 2na  z-hat
  double *zt = zh + 2 * na;         // 2Ma
  double *red = zt + 2 * Ma;        // 16

  const double2 *phi = reinterpret_cast<const double2 *>(bt.phi_d) + ln.base;
  const int T = ln.T;
  // 1. physical wave function of the stage input (FP64), sigma at stage 0
  double sigma;
  if (stage == 0) {
    zd acc = mk<double>(0.0, 0.0);
    for (int s = tid; s < n; s += nt) {
      const double2 v = phi[(long)s * T];
      acc.x += v.x * v.x + v.y * v.y;
    }
    acc = block_sum(acc, red);
    sigma = normed && !bt.sigma_raw ? 1.0 / sqrt(acc.x) : 1.0;
    for (int s = tid; s < n; s += nt) {
      const double2 v = phi[(long)s * T];
      psi[2 * s] = sigma * v.x;
      psi[2 * s + 1] = sigma * v.y;
      if (store_slot >= 0 && bt.psi_store_d) {
        double *dst = bt.psi_store_d + 2 * ((long)b * bt.store_stride + store_slot) * n;
        dst[2 * s] = sigma * v.x;
        dst[2 * s + 1] = sigma * v.y;
      }
    }
    if (tid == 0) {
      coef[L.sig] = sigma;
      bt.sigT_d[b] = sigma;
    }
  } else {
    sigma = eval ? 1.0 : coef[L.sig];
    if (eval && tid == 0) {
      coef[L.sig] = 1.0;
      bt.sigT_d[b] = 1.0;
    }
    const cx<R> *y = yin + ln.base;
    for (int s = tid; s < n; s += nt) {
      psi[2 * s] = (double)y[(long)s * T].x;
      psi[2 * s + 1] = (double)y[(long)s * T].y;
    }
  }
  __syncthreads();
  zd nrm = mk<double>(0.0, 0.0);
  for (int s = tid; s < n; s += nt) nrm.x += psi[2 * s] * psi[2 * s] + psi[2 * s + 1] * psi[2 * s + 1];
  nrm = block_sum(nrm, red);
  const bool has_lt = op.lt_d != nullptr;

  // 2. <L_j> (and <L_j^2> for the LTC) over active L-operators
  for (int j = tid; j < na; j += nt) {
    zd a = mk<double>(0.0, 0.0);
    for (int q = op.la_ptr_d[j]; q < op.la_ptr_d[j + 1]; ++q) {
      const int r0 = op.la_row_d[q], c0 = op.la_col_d[q];
      const zd v = cmul(cconj(ld2(psi, r0)), cmul(ld2(op.la_val_d, q), ld2(psi, c0)));
      a = cadd(a, v);
    }
    avg[2 * j] = a.x / nrm.x;
    avg[2 * j + 1] = a.y / nrm.x;
    zd a2 = mk<double>(0.0, 0.0);
    if (has_lt) {
      for (int q = op.lq_ptr_d[j]; q < op.lq_ptr_d[j + 1]; ++q) {
        const int r0 = op.lq_row_d[q], c0 = op.lq_col_d[q];
        a2 = cadd(a2, cmul(cconj(ld2(psi, r0)), cmul(ld2(op.lq_val_d, q), ld2(psi, c0))));
      }
    }
    avq[2 * j] = a2.x / nrm.x;
    avq[2 * j + 1] = a2.y / nrm.x;
  }
  __syncthreads();

  // 3. noise memory: stage value, derivative, RK4 accumulation
  //    (eom/eom_functions.py:65-160, integrator/integrator_rk.py:59-73)
  const double h = bt.dt_over_hbar;
  double *kz = ws;
  double *zacc = ws + 2 * Ma;
  for (int m = tid; m < Ma; m += nt) {
    zd z = ld2(zmem, m);
    if (nonlin && stage > 0) {
      const double c = c_rk[stage] * h;
      z.x += c * kz[2 * m];
      z.y += c * kz[2 * m + 1];
    }
    zt[2 * m] = z.x;
    zt[2 * m + 1] = z.y;
    if (!nonlin) continue;
    const int la = op.abs_lact_d[m];
    const zd g = ld2(op.abs_g_d, m), w = ld2(op.abs_w_d, m);
    zd d = mk<double>(0.0, 0.0);
    if (la >= -1) {
      const zd lav = la >= 0 ? ld2(avg, la) : mk<double>(0.0, 0.0);
      d = csub(cmul(lav, cconj(g)), cmul(cconj(w), z));
    } else if (z.x > 1e-8 || (z.x == 1e-8 && z.y > 0.0)) {
      d = mk<double>(0.0, 0.0);
      d = csub(d, cmul(cconj(w), z));
    }
    if (eval) continue;
    kz[2 * m] = d.x;
    kz[2 * m + 1] = d.y;
    if (stage == 0) {
      zacc[2 * m] = d.x;
      zacc[2 * m + 1] = d.y;
    } else if (stage < 3) {
      zacc[2 * m] += 2.0 * d.x;
      zacc[2 * m + 1] += 2.0 * d.y;
    } else {
      const double c = h / 6.0;
      zmem[2 * m] += c * (zacc[2 * m] + d.x);
      zmem[2 * m + 1] += c * (zacc[2 * m + 1] + d.y);
    }
  }
  __syncthreads();

  // 4. z-hat and the noise coefficient of each active L-operator
  // noise values are complex64, or complex128 (interpolated or
  // user-supplied complex128 noise, used unrounded as the reference uses it)
  const bool z128 = bt.noise_c128 != 0;
  auto znoise = [&](const void *base, long idx) -> zd {
    if (z128) {
      const double2 v = reinterpret_cast<const double2 *>(base)[idx];
      return mk<double>(v.x, v.y);
    }
    const float2 v = reinterpret_cast<const float2 *>(base)[idx];
    return mk<double>((double)v.x, (double)v.y);
  };
  const long zoff = (long)b * na * bt.n_t;
  for (int j = tid; j < na; j += nt) {
    const long zr = op.act_zrow_d ? op.act_zrow_d[j] : j;
    const zd v1 = znoise(bt.z1_d, zoff + zr * bt.n_t + tidx);
    zd zz = mk<double>(v1.x, -v1.y);
    if (nonlin) {
      for (int q = op.am_ptr_d[j]; q < op.am_ptr_d[j + 1]; ++q) {
        const int ma = op.mode_abs_d[op.am_mode_d[q]];
        zz.x += zt[2 * ma];
        zz.y += zt[2 * ma + 1];
      }
    }
    zh[2 * j] = zz.x;
    zh[2 * j + 1] = zz.y;
    zd c = zz;
    if (bt.z2_d) {
      const zd v2 = znoise(bt.z2_d, zoff + zr * bt.n_t + tidx);
      if (nonlin) {
        c.x += v2.y;
        c.y -= v2.x;
      } else if (z128) {
        c.x = v1.x + v2.y;
        c.y = -v1.y - v2.x;
      } else {
        // LINEAR: conj(z1) - i z2 of two complex64 noise values is itself
        // complex64, rounded in single precision (eom/hops_eom.py:519,534-536)
        c.x = (double)((float)v1.x + (float)v2.y);
        c.y = (double)(-(float)v1.y - (float)v2.x);
      }
    }
    cj[2 * j] = c.x;
    cj[2 * j + 1] = c.y;
  }
  __syncthreads();

  // 5. normalisation correction (eom/eom_functions.py:163-213, :258-284)
  zd nc = mk<double>(0.0, 0.0);
  if (normed) {
    zd part = mk<double>(0.0, 0.0);
    for (int j = tid; j < na; j += nt) part = cadd(part, cmul(ld2(zh, j), ld2(avg, j)));
    for (int f = tid; f < op.n_first; f += nt) {
      const int ia = op.first_aux_d[f], l = op.first_lact_d[f], m = op.first_mode_d[f];
      const zd gw = cdiv(ld2(op.mode_g_d, m), ld2(op.mode_w_d, m));
      // rows of the first-level auxiliary in the stage input
      zd lterm = mk<double>(0.0, 0.0), oterm = mk<double>(0.0, 0.0);
      for (int q = op.la_ptr_d[l]; q < op.la_ptr_d[l + 1]; ++q) {
        const int r0 = op.la_row_d[q], c0 = op.la_col_d[q];
        zd pe;
        if (stage == 0) {
          const double2 v = phi[((long)ia * n + c0) * T];
          pe = mk<double>(sigma * v.x, sigma * v.y);
        } else {
          pe = ccast<double>(yin[ln.base + ((long)ia * n + c0) * T]);
        }
        lterm = cadd(lterm, cmul(cconj(ld2(psi, r0)), cmul(ld2(op.la_val_d, q), cmul(gw, pe))));
      }
      for (int s = 0; s < n; ++s) {
        zd pe;
        if (stage == 0) {
          const double2 v = phi[((long)ia * n + s) * T];
          pe = mk<double>(sigma * v.x, sigma * v.y);
        } else {
          pe = ccast<double>(yin[ln.base + ((long)ia * n + s) * T]);
        }
        oterm = cadd(oterm, cmul(cconj(ld2(psi, s)), cmul(gw, pe)));
      }
      part = csub(part, lterm);
      part = cadd(part, cmul(oterm, ld2(avg, l)));
    }
    part = block_sum(part, red);
    nc = mk<double>(part.x, 0.0);
    if (has_lt) {
      zd q = mk<double>(0.0, 0.0);
      for (int j = tid; j < na; j += nt) {
        const zd a = ld2(avg, j);
        const zd t = csub(cscale(cmul(a, a), 2.0), ld2(avq, j));
        q = cadd(q, cscale(t, op.lt_d[2 * j]));
      }
      q = block_sum(q, red);
      nc = cadd(nc, q);
    }
  }

  // 6. system-block coefficients, Z2 coefficients (FP64 per trajectory for
  //    this kernel, and in stage precision trajectory-fast for k_stage)
  cx<R> *cT = reinterpret_cast<cx<R> *>(bt.coefT_d);
  const long cs = bt.coefT_stride;
  for (int p = tid; p < op.n_pat; p += nt) {
    zd a = ld2(op.pat_h_d, p), a0 = a;
    for (int e = op.pc_ptr_d[p]; e < op.pc_ptr_d[p + 1]; ++e) {
      const int j = op.pc_act_d[e];
      const zd v = ld2(op.pc_val_d, e);
      if (op.pc_kind_d[e] == 0) {
        zd c = ld2(cj, j);
        zd c0 = c;
        if (nonlin && has_lt) {
          const zd lt = ld2(op.lt_d, j), av = ld2(avg, j);
          c = cadd(c, cmul(cconj(lt), av));
          c0 = cadd(c, cmul(lt, av));
        }
        a = cadd(a, cmul(c, v));
        a0 = cadd(a0, cmul(c0, v));
      } else {
        a0 = csub(a0, cmul(ld2(op.lt_d, j), v));
      }
    }
    coef[L.A + 2 * p] = a.x;
    coef[L.A + 2 * p + 1] = a.y;
    coef[L.A0 + 2 * p] = a0.x;
    coef[L.A0 + 2 * p + 1] = a0.y;
    cT[(long)p * cs + b] = ccast<R>(a);
    cT[(long)(op.n_pat + p) * cs + b] = ccast<R>(a0);
  }
  if (nonlin) {
    for (int m = tid; m < op.n_mode; m += nt) {
      const int la = op.mode_lact_d[m];
      zd z = mk<double>(0.0, 0.0);
      if (la >= 0) z = cmul(cdiv(ld2(op.mode_g_d, m), ld2(op.mode_w_d, m)), cconj(ld2(avg, la)));
      coef[L.zc + 2 * m] = z.x;
      coef[L.zc + 2 * m + 1] = z.y;
      cT[(long)(2 * op.n_pat + m) * cs + b] = ccast<R>(z);
    }
  }
  if (tid == 0) {
    coef[L.nc] = nc.x;
    coef[L.nc + 1] = nc.y;
    cT[(long)(2 * op.n_pat + op.n_mode) * cs + b] = ccast<R>(nc);
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_prologue(const hops_op_t op, const hops_batch_t bt,
                                                  int stage, int tidx, const cx<R> *__restrict__ yin,
                                                  int store_slot, const int *__restrict__ step_ctr) {
  if (step_ctr) {  // replayed graph: noise index and store slot advance by the counter
    tidx += step_ctr[0];
    if (store_slot >= 0) store_slot += step_ctr[1];
  }
  prologue_body<R>(op, bt, blockIdx.x, stage, tidx, yin, store_slot);
}

// The prologues of all four stages of one LINEAR step in one launch.  LINEAR
// stage coefficients depend on the noise only, not on the stage input
// (eom/hops_eom.py:443-542), so workgroup (b, stage) can run before the step:
// stage 0 writes the batch's own tables (and the psi store, sigma), stages
// 1..3 write the tables coefT_s / coef_s + (stage - 1) * blk that their
// stage kernels then read.
template <typename R>
__global__ __launch_bounds__(256) void k_prologue_linear4(const hops_op_t op, hops_batch_t bt, int tidx0, int ns,
                                                         const cx<R> *__restrict__ yin, int store_slot,
                                                         char *coefT_s, size_t coefT_blk, double *coef_s,
                                                         size_t coef_blk, const int *__restrict__ step_ctr) {
  const int stage = blockIdx.y;
  if (step_ctr) {
    tidx0 += step_ctr[0];
    if (store_slot >= 0) store_slot += step_ctr[1];
  }
  const int off = stage == 0 ? 0 : (stage == 3 ? 2 * ns : ns);  // noise index map [0, 1, 1, 2]
  if (stage > 0) {
    bt.coefT_d = coefT_s + (size_t)(stage - 1) * coefT_blk;
    bt.coef_d = coef_s + (size_t)(stage - 1) * coef_blk;
  }
  prologue_body<R>(op, bt, blockIdx.x, stage, tidx0 + off, yin, stage == 0 ? store_slot : -1);
}

// Derivative of row r = (k, s) of trajectory b fused with its RK4 stage
// update (eom/hops_eom.py:386-426, integrator/integrator_rk.py:53-73).  The
// ELL row of the super-operator comes in ecol/ev: registers when WR > 0,
// otherwise the column-major ELL arrays in HBM.
template <typename R, int WR>
__device__ __forceinline__ void stage_row(const hops_op_t &op, const hops_batch_t &bt, int stage,
                                          const cx<R> *__restrict__ yin, cx<R> *__restrict__ yout, int b,
                                          long r, int k, int s, cx<R> k0, const int *ecol, const cx<R> *ev) {
#pragma clang fp contract(off)
  const int n = op.n_state;
  const long nr = (long)n * op.n_hier;
  const Lane ln = lane_of(bt, nr, b);
  const int T = ln.T;
  const cx<R> *Y = yin + ln.base;
  const cx<R> *cT = reinterpret_cast<const cx<R> *>(bt.coefT_d) + b;
  const long cs = bt.coefT_stride;
  const double sigma = bt.sigT_d[b];

  const cx<R> yr = Y[r * T];
  cx<R> d = cmul(k0, yr);
  // dense system block: (-iH) Y from the matrix cores (k_dense_h)
  if (bt.hy_d) d = cadd(d, reinterpret_cast<const cx<R> *>(bt.hy_d)[ln.base + r * T]);
  // system block: -iH (when not dense) + noise / LTC terms
  const cx<R> *A = cT + (k == 0 ? (long)op.n_pat * cs : 0);
  const long rowk = (long)k * n;
  for (int p = op.pat_ptr_d[s]; p < op.pat_ptr_d[s + 1]; ++p)
    cfma(d, A[p * cs], Y[(rowk + op.pat_col_d[p]) * T]);
  // hierarchy super-operator K2_kp1 + K2_km1
  if (WR > 0) {
#pragma unroll
    for (int e = 0; e < WR; ++e)
      if (ecol[e] >= 0) cfma(d, ev[e], Y[(long)ecol[e] * T]);
  } else {
    const cx<R> *evm = reinterpret_cast<const cx<R> *>(op.ell_val_d);
    for (int e = 0; e < op.ell_w; ++e) {
      const int c = op.ell_col_d[(long)e * nr + r];
      if (c < 0) break;
      cfma(d, evm[(long)e * nr + r], Y[(long)c * T]);
    }
  }
  if (bt.eom != HOPS_EOM_LINEAR) {
    const cx<R> *zc = cT + 2L * op.n_pat * cs;
    for (int u = op.up_ptr_d[k]; u < op.up_ptr_d[k + 1]; ++u)
      cfma(d, zc[(long)op.up_mode_d[u] * cs], Y[((long)op.up_aux_d[u] * n + s) * T]);
    if (bt.eom == HOPS_EOM_NORMALIZED_NONLINEAR) {
      const cx<R> nc = cT[(2L * op.n_pat + op.n_mode) * cs];
      d = csub(d, cmul(nc, yr));
    }
  }
  if (stage == 0) d = cscale(d, (R)sigma);

  const long gi = ln.base + r * T;
  if (stage < 0) {
    yout[gi] = d;
    return;
  }
  const double h = bt.dt_over_hbar;
  cx<R> *acc = reinterpret_cast<cx<R> *>(bt.acc_d);
  double2 *phi = reinterpret_cast<double2 *>(bt.phi_d);
  if (stage < 3) {
    cx<R> base;
    cx<R> a_in = mk<R>((R)0, (R)0);
    if (stage > 0) a_in = acc[gi];
    if (stage == 1) {
      // the stage base sigma * y0 from this row's stage-1 input and k0:
      // y1 - c0 k0, one rounding of y1 away from it, without reading y0
      const R c0 = (R)(0.5 * h);
      base = mk<R>(yr.x - c0 * a_in.x, yr.y - c0 * a_in.y);
    } else if (bt.precision == HOPS_PREC_MIXED) {
      base = cscale(reinterpret_cast<const cx<R> *>(bt.y0_d)[gi], (R)sigma);
    } else {
      const double2 p = phi[gi];
      base = mk<R>((R)(sigma * p.x), (R)(sigma * p.y));
    }
    const R c = (R)((stage == 2 ? 1.0 : 0.5) * h);
    yout[gi] = mk<R>(base.x + c * d.x, base.y + c * d.y);
    if (stage == 0) {
      acc[gi] = d;
    } else {
      cx<R> a = a_in;
      a.x += (R)2 * d.x;
      a.y += (R)2 * d.y;
      acc[gi] = a;
    }
  } else {
    const cx<R> a = acc[gi];
    const double2 p = phi[gi];
    const double c = h / 6.0;
    double2 q;
    q.x = sigma * p.x + c * ((double)a.x + (double)d.x);
    q.y = sigma * p.y + c * ((double)a.y + (double)d.y);
    phi[gi] = q;
    if (bt.precision == HOPS_PREC_MIXED)
      reinterpret_cast<cx<float> *>(bt.y0_d)[gi] = mk<float>((float)q.x, (float)q.y);
  }
}

// Streaming form: one thread per (row, lane); workgroups from the tile list,
// a tile being 256 / T rows of one trajectory group.
template <typename R>
__global__ __launch_bounds__(256) void k_stage(const hops_op_t op, const hops_batch_t bt, int stage,
                                               const cx<R> *__restrict__ yin,
                                               cx<R> *__restrict__ yout) {
  const int t = bt.tile_d[blockIdx.x];
  if (t < 0) return;
  const int g = t >> 16, tile = t & 0xffff;
  const int lane = threadIdx.x & ((1 << bt.lane_shift) - 1);
  const int b = (g << bt.lane_shift) + lane;
  if (b >= bt.B) return;
  const int n = op.n_state;
  const long nr = (long)n * op.n_hier;
  const long r = (long)tile * (bt.tile_rows >> bt.lane_shift) + (threadIdx.x >> bt.lane_shift);
  if (r >= nr) return;
  const int k = (int)(r / n), s = (int)(r - (long)k * n);
  stage_row<R, 0>(op, bt, stage, yin, yout, b, r, k, s, ccast<R>(ld2(op.k0_d, k)), nullptr, nullptr);
}

// Dense system block on the matrix cores: HY = (-iH) Y for every auxiliary
// and trajectory lane of a stage input (the reference's (-1j*H) @ Phi,
// eom/hops_eom.py:389-391, 526).  For each trajectory group the product is
// the GEMM C[s, j] = sum_s' A[s, s'] Y[(k, s'), t] over the columns
// j = k * T + t.  A wave computes one 16 x 16 tile of C with the 16x16x4 MFMA
// (f32 for complex64 stages, f64 for complex128), the complex product as four
// real ones: Cr += Ar Yr + (-Ai) Yi, Ci += Ar Yi + Ai Yr.  Lane l supplies
// A[i0 + l % 16][kk + l / 16] and Y[kk + l / 16][j0 + l % 16]; element v of its
// accumulator is C[i0 + 4 (l / 16) + v][j0 + l % 16] for f32 and
// C[i0 + l / 16 + 4 v][j0 + l % 16] for f64 (both measured on gfx950).  A
// workgroup holds four waves, four row tiles of one column tile (the same Y
// fragments, from L1).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f32x4 mfma16(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mfma16(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
template <typename R> struct Acc4;
template <> struct Acc4<float> {
  typedef f32x4 T;
  // accumulator row of element v of a lane in row quarter q = lane / 16
  static __device__ __forceinline__ int row(int q, int v) { return 4 * q + v; }
};
template <> struct Acc4<double> {
  typedef f64x4 T;
  static __device__ __forceinline__ int row(int q, int v) { return q + 4 * v; }
};

template <typename R>
__global__ __launch_bounds__(256) void k_dense_h(const hops_op_t op, const hops_batch_t bt,
                                                 const cx<R> *__restrict__ yin, cx<R> *__restrict__ hy) {
  typedef typename Acc4<R>::T V4;
  const int n = op.n_state, hn = op.hd_n, ls = bt.lane_shift;
  const int T = 1 << ls;
  const long nr = (long)n * op.n_hier;
  const long ncol = (long)op.n_hier << ls;
  const long n_ct = (ncol + 15) >> 4;
  const int n_rt = hn >> 4, n_rq = (n_rt + 3) >> 2;
  long blk = blockIdx.x;
  const int rq = (int)(blk % n_rq);
  blk /= n_rq;
  const long ct = blk % n_ct;
  const long g = blk / n_ct;
  const int rt = rq * 4 + (int)(threadIdx.x >> 6);
  if (rt >= n_rt) return;  // whole waves
  const int l = threadIdx.x & 63;
  const int i0 = rt << 4;
  const long j = (ct << 4) + (l & 15);
  const bool jv = j < ncol;
  const long kaux = jv ? (j >> ls) : 0;
  const int t = (int)(j & (T - 1));
  const long cbase = (g * nr + kaux * n) * T + t;  // Y[(kaux, s'), t] at cbase + s' * T
  const R *Ar = reinterpret_cast<const R *>(op.hd_d);
  const R *Ai = Ar + (long)hn * hn;
  const long arow = (long)(i0 + (l & 15)) * hn;
  const int kq = l >> 4;
  V4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
  for (int kk = 0; kk < hn; kk += 4) {
    const int kc = kk + kq;
    const R ar = Ar[arow + kc], ai = Ai[arow + kc];
    cx<R> y = mk<R>((R)0, (R)0);
    if (jv && kc < n) y = yin[cbase + (long)kc * T];
    cr = mfma16(ar, y.x, cr);
    cr = mfma16(-ai, y.y, cr);
    ci = mfma16(ar, y.y, ci);
    ci = mfma16(ai, y.x, ci);
  }
  if (!jv) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int s = i0 + Acc4<R>::row(kq, v);
    if (s < n) hy[cbase + (long)s * T] = mk<R>(cr[v], ci[v]);
  }
}

// grid of k_dense_h for a batch of G trajectory groups
unsigned dense_grid(const hops_op_t &op, int G, int lane_shift) {
  const long ncol = (long)op.n_hier << lane_shift;
  const long n_ct = (ncol + 15) >> 4;
  const int n_rq = ((op.hd_n >> 4) + 3) >> 2;
  return (unsigned)((long)G * n_ct * n_rq);
}

// LDS-staged form of k_dense_h: one workgroup per 16-column tile computes
// every row tile of it.  The tile's Y block (hd_n rows x 16 columns, zero
// rows past n_state) is read from HBM once, coalesced, into LDS (column c at
// ys[c * hs + s'], hs = hd_n + 2), and the four waves take the row tiles
// rt = w, w + 4, ... two at a time (independent accumulators).  Within each
// 16-deep slice kb of the sum, MFMA step v pairs lane (q, r)'s A[i0 + r][kb +
// 4q + v] with Y[kb + 4q + v][j0 + r]: every product enters exactly once, and
// lane (q, r) reads its four A values of the slice as one 16-byte (f32) or
// 32-byte (f64) load and its four Y values from consecutive LDS words.
template <typename R> struct Vec4;
template <> struct Vec4<float> { typedef float T __attribute__((ext_vector_type(4))); };
template <> struct Vec4<double> { typedef double T __attribute__((ext_vector_type(4))); };

template <typename R>
__global__ __launch_bounds__(256) void k_dense_lds(const hops_op_t op, const hops_batch_t bt,
                                                   const cx<R> *__restrict__ yin, cx<R> *__restrict__ hy,
                                                   int hs) {
  typedef typename Acc4<R>::T V4;
  typedef typename Vec4<R>::T A4;
  extern __shared__ __align__(16) unsigned char dense_lds[];
  cx<R> *ys = reinterpret_cast<cx<R> *>(dense_lds);
  const int n = op.n_state, hn = op.hd_n, ls = bt.lane_shift;
  const int T = 1 << ls, Tc = T < 16 ? T : 16;
  const long nr = (long)n * op.n_hier;
  const long ncol = (long)op.n_hier << ls;
  const long n_ct = (ncol + 15) >> 4;
  const long ct = (long)blockIdx.x % n_ct, g = (long)blockIdx.x / n_ct;
  const long j0 = ct << 4;
  // the Y block: e runs over (column group, row, lane) with the lane fastest,
  // the order of the pool (aux-major rows, T lanes interleaved)
  for (int e = threadIdx.x; e < 16 * hn; e += 256) {
    const int grp = e / (hn * Tc), rem = e - grp * hn * Tc;
    const int sp = rem / Tc, tt = rem - sp * Tc;
    const int c = grp * Tc + tt;
    const long j = j0 + c;
    cx<R> v = mk<R>((R)0, (R)0);
    if (sp < n && j < ncol) v = yin[(g * nr + (j >> ls) * n + sp) * T + (j & (T - 1))];
    ys[c * hs + sp] = v;
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, q = l >> 4, r = l & 15;
  const int n_rt = hn >> 4;
  const R *Ar = reinterpret_cast<const R *>(op.hd_d);
  const R *Ai = Ar + (long)hn * hn;
  const long j = j0 + r;
  const bool jv = j < ncol;
  const long cbase = jv ? (g * nr + (j >> ls) * n) * T + (j & (T - 1)) : 0;
  const cx<R> *yc = ys + r * hs + 4 * q;
  for (int ra = w; ra < n_rt; ra += 8) {
    const int rb = ra + 4;
    const bool two = rb < n_rt;
    const A4 *pa = reinterpret_cast<const A4 *>(Ar + (long)(ra * 16 + r) * hn + 4 * q);
    const A4 *pai = reinterpret_cast<const A4 *>(Ai + (long)(ra * 16 + r) * hn + 4 * q);
    const A4 *pb = reinterpret_cast<const A4 *>(Ar + (long)((two ? rb : ra) * 16 + r) * hn + 4 * q);
    const A4 *pbi = reinterpret_cast<const A4 *>(Ai + (long)((two ? rb : ra) * 16 + r) * hn + 4 * q);
    V4 cra = {0, 0, 0, 0}, cia = {0, 0, 0, 0}, crb = {0, 0, 0, 0}, cib = {0, 0, 0, 0};
    for (int kb = 0; kb < hn; kb += 16) {
      const A4 a = pa[kb >> 2], ai = pai[kb >> 2];
      A4 b = a, bi = ai;
      if (two) {
        b = pb[kb >> 2];
        bi = pbi[kb >> 2];
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const cx<R> y = yc[kb + v];
        cra = mfma16(a[v], y.x, cra);
        cra = mfma16(-ai[v], y.y, cra);
        cia = mfma16(a[v], y.y, cia);
        cia = mfma16(ai[v], y.x, cia);
        if (two) {
          crb = mfma16(b[v], y.x, crb);
          crb = mfma16(-bi[v], y.y, crb);
          cib = mfma16(b[v], y.y, cib);
          cib = mfma16(bi[v], y.x, cib);
        }
      }
    }
    if (jv) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int sa = ra * 16 + Acc4<R>::row(q, v);
        if (sa < n) hy[cbase + (long)sa * T] = mk<R>(cra[v], cia[v]);
        const int sb = rb * 16 + Acc4<R>::row(q, v);
        if (two && sb < n) hy[cbase + (long)sb * T] = mk<R>(crb[v], cib[v]);
      }
    }
  }
}

// (-iH) Y of G trajectory groups: the LDS-staged form when its Y block fits
// in 64 KiB of LDS, else the plain form
template <typename R>
int launch_dense_h(const hops_op_t &op, const hops_batch_t &bt, int G, const cx<R> *yin, cx<R> *hy,
                   hipStream_t st) {
  const int hs = op.hd_n + 2;
  const size_t lds = (size_t)16 * hs * sizeof(cx<R>);
  if (lds <= 64 * 1024) {
    const long n_ct = (((long)op.n_hier << bt.lane_shift) + 15) >> 4;
    k_dense_lds<R><<<(unsigned)(G * n_ct), 256, lds, st>>>(op, bt, yin, hy, hs);
  } else {
    k_dense_h<R><<<dense_grid(op, G, bt.lane_shift), 256, 0, st>>>(op, bt, yin, hy);
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

// Resident form for small bases: one workgroup per trajectory runs every
// RK4 step of the call, the four prologues and stages of a step separated by
// workgroup barriers (a trajectory's rows only read rows of the same
// trajectory).  Each stage is the same stage_row per (row, lane) as k_stage,
// so results are bit-identical to the launched forms; a call costs one
// launch instead of five or eight per step.
template <typename R>
__global__ __launch_bounds__(256) void k_steps_resident(const hops_op_t op, const hops_batch_t bt, int it0,
                                                        int n_steps, int store0, const cx<R> *__restrict__ y_first,
                                                        cx<R> *__restrict__ ya, cx<R> *__restrict__ yb) {
  const int b = blockIdx.x;
  if (b >= bt.B) return;
  const int n = op.n_state, ns = bt.noise_step;
  const long nr = (long)n * op.n_hier;
  const int toff[4] = {0, ns, ns, 2 * ns};
  for (int step = 0; step < n_steps; ++step) {
    const int itb = it0 + step * 2 * ns;
    for (int i = 0; i < 4; ++i) {
      const cx<R> *yin = i == 0 ? y_first : (i == 2 ? yb : ya);
      cx<R> *yout = i == 1 ? yb : ya;
      const int slot = (i == 0 && step > 0 && bt.psi_store_d) ? store0 + step - 1 : -1;
      prologue_body<R>(op, bt, b, i, itb + toff[i], yin, slot);
      __syncthreads();
      for (long r = threadIdx.x; r < nr; r += blockDim.x) {
        const int k = (int)(r / n), s = (int)(r - (long)k * n);
        stage_row<R, 0>(op, bt, i, yin, yout, b, r, k, s, ccast<R>(ld2(op.k0_d, k)), nullptr, nullptr);
      }
      __syncthreads();
    }
  }
}

// Element i of an array addressed as a uniform base pointer plus a 32-bit
// byte offset, which the compiler issues as a global load/store with a scalar
// base and one offset register (the pools of the stationary kernel are below
// 4 GiB; the launch checks it).
template <typename T>
__device__ __forceinline__ T *el(T *base, uint32_t i) {
  return reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (uint64_t)(i * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ const T *el(const T *base, uint32_t i) {
  return reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (uint64_t)(i * (uint32_t)sizeof(T)));
}

// V consecutive elements starting at p (p aligned to V * sizeof(T)) as one
// 8-byte access or a run of 16-byte accesses, optionally nontemporal.
template <int NT, int V, typename T>
__device__ __forceinline__ void ldv(const T *p, T (&o)[V]) {
  constexpr int NB = V * (int)sizeof(T);
  if constexpr (NB == 8) {
    o[0] = ld_s<NT>(p);
  } else {
    static_assert(NB % 16 == 0, "16-byte runs");
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    const u4 *q = reinterpret_cast<const u4 *>(p);
#pragma unroll
    for (int i = 0; i < NB / 16; ++i) {
      u4 u;
      if constexpr (NT != 0) u = __builtin_nontemporal_load(q + i);
      else u = q[i];
      __builtin_memcpy(reinterpret_cast<char *>(o) + 16 * i, &u, 16);
    }
  }
}
template <int NT, int V, typename T>
__device__ __forceinline__ void stv(T *p, const T (&v)[V]) {
  constexpr int NB = V * (int)sizeof(T);
  if constexpr (NB == 8) {
    st_s<NT>(p, v[0]);
  } else {
    static_assert(NB % 16 == 0, "16-byte runs");
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    u4 *q = reinterpret_cast<u4 *>(p);
#pragma unroll
    for (int i = 0; i < NB / 16; ++i) {
      u4 u;
      __builtin_memcpy(&u, reinterpret_cast<const char *>(v) + 16 * i, 16);
      if constexpr (NT != 0) __builtin_nontemporal_store(u, q + i);
      else q[i] = u;
    }
  }
}

// Row update of V adjacent trajectory lanes b .. b+V-1 of one row (b a
// multiple of V), with the row's pattern and ELL entries in registers and
// every load issued before the arithmetic.  The lanes of a row are contiguous
// in every pool, so each operand of the V lanes is one wide access: V = 2
// halves the vector-memory instructions per element, V = 4 quarters them.
// Each lane performs exactly the operation sequence of stage_row.
// pidx[p] < 0 and ecol[e] < 0 mark unused slots.  NT >= 1: the accumulator
// and the carry are read and written with the nontemporal hint; NT >= 2: the
// stage output and the complex64 state copy are stored with it too; NT >= 3:
// the row's own aux block is read with it; NT = 4: the ELL neighbours too.
//
// LDS = true (k_stage_lds): the row's own value comes in yr_pre, and its
// pattern values and coefficients are read from LDS: ly points at lane b of
// the row's aux block (element s' of lane b + v at ly[s' T + v]), lc at lane
// b of the coefficient set of the row (coefficient p at lc[p T + v]).
template <typename R, int WR, int PW, int NT, int V, bool LDS = false, bool HY = false>
__device__ __forceinline__ void stage_row_reg(const hops_op_t &op, const hops_batch_t &bt, int stage,
                                              const cx<R> *__restrict__ yin, cx<R> *__restrict__ yout,
                                              int b, long r_, int k, cx<R> k0, const int *ecol,
                                              const cx<R> *ev, const int *pidx, const int *pcol,
                                              const cx<R> *ly = nullptr, const cx<R> *lc = nullptr,
                                              const cx<R> *yr_pre = nullptr) {
#pragma clang fp contract(off)
  const uint32_t n = (uint32_t)op.n_state;
  const uint32_t nr = n * (uint32_t)op.n_hier;
  const uint32_t r = (uint32_t)r_;
  const uint32_t ls = (uint32_t)bt.lane_shift, T = 1u << ls;
  const uint32_t gb = ((uint32_t)b >> ls) * nr * T + ((uint32_t)b & (T - 1));  // lane base
  const uint32_t rowk = (uint32_t)k * n;
  const uint32_t gi = gb + r * T;
  const cx<R> *cT = reinterpret_cast<const cx<R> *>(bt.coefT_d);
  const uint32_t cs = (uint32_t)bt.coefT_stride;
  cx<R> *acc = reinterpret_cast<cx<R> *>(bt.acc_d);
  double2 *phi = reinterpret_cast<double2 *>(bt.phi_d);
  const cx<R> zero = mk<R>((R)0, (R)0);
  // loads
  constexpr int NTO = NT >= 3 ? 1 : 0;  // the row's own aux block
  constexpr int NTG = NT >= 4 ? 1 : 0;  // the ELL neighbours
  cx<R> yr[V];
  if constexpr (LDS) {
#pragma unroll
    for (int v = 0; v < V; ++v) yr[v] = yr_pre[v];
  } else {
    ldv<NTO, V>(el(yin, gi), yr);
  }
  cx<R> yp[PW][V], ye[WR][V], ap[PW][V];
  const uint32_t a0 = (uint32_t)b + (k == 0 ? (uint32_t)op.n_pat * cs : 0u);
#pragma unroll
  for (int p = 0; p < PW; ++p) {
    if (pidx[p] >= 0) {
      if constexpr (LDS) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
          yp[p][v] = ly[pcol[p] * T + v];
          ap[p][v] = lc[pidx[p] * T + v];
        }
      } else {
        ldv<NTO, V>(el(yin, gb + (rowk + (uint32_t)pcol[p]) * T), yp[p]);
        ldv<0, V>(el(cT, a0 + (uint32_t)pidx[p] * cs), ap[p]);
      }
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) yp[p][v] = ap[p][v] = zero;
    }
  }
#pragma unroll
  for (int e = 0; e < WR; ++e) {
    if (ecol[e] >= 0) {
      ldv<NTG, V>(el(yin, gb + (uint32_t)ecol[e] * T), ye[e]);
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) ye[e][v] = zero;
    }
  }
  // the dense system block's (-iH) Y of this row (k_dense_lds / k_dense_h)
  cx<R> hyv[V];
  if constexpr (HY) ldv<NT >= 1 ? 1 : 0, V>(el(reinterpret_cast<const cx<R> *>(bt.hy_d), gi), hyv);
  double sigma[V];
  ldv<0, V>(bt.sigT_d + b, sigma);
  cx<R> base[V], a_in[V];
  double2 ph[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    base[v] = a_in[v] = zero;
    ph[v] = make_double2(0.0, 0.0);
  }
  if (stage == 1) {
    // the stage base comes from this row's input and k0 (below)
  } else if (stage < 3) {
    if (bt.precision == HOPS_PREC_MIXED) {
      ldv<NT, V>(el(reinterpret_cast<const cx<R> *>(bt.y0_d), gi), base);
    } else {
      ldv<NT, V>(el(phi, gi), ph);
    }
  } else {
    ldv<NT, V>(el(phi, gi), ph);
  }
  if (stage > 0) ldv<NT, V>(el(acc, gi), a_in);
  // arithmetic, lane by lane
  cx<R> d[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    d[v] = cmul(k0, yr[v]);
    if constexpr (HY) d[v] = cadd(d[v], hyv[v]);
#pragma unroll
    for (int p = 0; p < PW; ++p) cfma(d[v], ap[p][v], yp[p][v]);
#pragma unroll
    for (int e = 0; e < WR; ++e) cfma(d[v], ev[e], ye[e][v]);
  }
  if (bt.eom != HOPS_EOM_LINEAR) {
    const uint32_t zc0 = (uint32_t)b + 2u * (uint32_t)op.n_pat * cs;
    for (int u = op.up_ptr_d[k]; u < op.up_ptr_d[k + 1]; ++u) {
      cx<R> zc[V], yu[V];
      ldv<0, V>(el(cT, zc0 + (uint32_t)op.up_mode_d[u] * cs), zc);
      ldv<0, V>(el(yin, gb + ((uint32_t)op.up_aux_d[u] * n + (r - rowk)) * T), yu);
#pragma unroll
      for (int v = 0; v < V; ++v) cfma(d[v], zc[v], yu[v]);
    }
    if (bt.eom == HOPS_EOM_NORMALIZED_NONLINEAR) {
      cx<R> nc[V];
      ldv<0, V>(el(cT, (uint32_t)b + (2u * (uint32_t)op.n_pat + (uint32_t)op.n_mode) * cs), nc);
#pragma unroll
      for (int v = 0; v < V; ++v) d[v] = csub(d[v], cmul(nc[v], yr[v]));
    }
  }
  const double h = bt.dt_over_hbar;
  if (stage < 3) {
    const R c = (R)((stage == 2 ? 1.0 : 0.5) * h);
    cx<R> o[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      if (stage == 0) d[v] = cscale(d[v], (R)sigma[v]);
      if (stage == 1) {
        // sigma * y0 = y1 - c0 k0, one rounding of y1 away, as in stage_row
        const R c0 = (R)(0.5 * h);
        base[v] = mk<R>(yr[v].x - c0 * a_in[v].x, yr[v].y - c0 * a_in[v].y);
      } else if (bt.precision == HOPS_PREC_MIXED) {
        base[v] = cscale(base[v], (R)sigma[v]);
      } else {
        base[v] = mk<R>((R)(sigma[v] * ph[v].x), (R)(sigma[v] * ph[v].y));
      }
      o[v] = mk<R>(base[v].x + c * d[v].x, base[v].y + c * d[v].y);
      if (stage == 0) {
        a_in[v] = d[v];
      } else {
        a_in[v].x += (R)2 * d[v].x;
        a_in[v].y += (R)2 * d[v].y;
      }
    }
    stv<NT >= 2 ? 1 : 0, V>(el(yout, gi), o);
    stv<NT, V>(el(acc, gi), a_in);
  } else {
    const double c = h / 6.0;
    double2 q[V];
    cx<float> q32[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      q[v].x = sigma[v] * ph[v].x + c * ((double)a_in[v].x + (double)d[v].x);
      q[v].y = sigma[v] * ph[v].y + c * ((double)a_in[v].y + (double)d[v].y);
      q32[v] = mk<float>((float)q[v].x, (float)q[v].y);
    }
    stv<NT, V>(el(phi, gi), q);
    if (bt.precision == HOPS_PREC_MIXED) stv<NT >= 2 ? 1 : 0, V>(el(reinterpret_cast<cx<float> *>(bt.y0_d), gi), q32);
  }
}

// Operator-stationary form for a batch that shares one operator: workgroup i
// owns a tile of 512 V / T rows (T / V threads per row, V lanes each), keeps
// its K0 and ELL rows in registers and walks the trajectory groups
// g = x + 8 (q + Q j) of XCD x = i % 8, so the operator is read once per tile
// and stage instead of once per trajectory, and every XCD works on the same
// few groups at a time (their rows stay in its L2).
template <typename R, int WR, int PW, int NT, int V, bool HY = false>
__global__ __launch_bounds__(512) void k_stage_shared(const hops_op_t op, const hops_batch_t bt, int stage,
                                                      const cx<R> *__restrict__ yin,
                                                      cx<R> *__restrict__ yout, int n_tiles, int Q) {
  const int x = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int q = slot / n_tiles;
  const int tile = bt.tile_order_d ? bt.tile_order_d[slot - q * n_tiles] : slot - q * n_tiles;
  const int n = op.n_state;
  const long nr = (long)n * op.n_hier;
  const int tps = bt.lane_shift - (V == 4 ? 2 : (V == 2 ? 1 : 0));  // log2 threads per row
  const int lane = (threadIdx.x & ((1 << tps) - 1)) * V;
  const long r = (long)tile * (512 >> tps) + (threadIdx.x >> tps);
  if (r >= nr) return;
  const int k = (int)(r / n), s = (int)(r - (long)k * n);
  int ecol[WR];
  cx<R> ev[WR];
  const cx<R> *evm = reinterpret_cast<const cx<R> *>(op.ell_val_d);
#pragma unroll
  for (int e = 0; e < WR; ++e) {
    if (e < op.ell_w) {
      ecol[e] = op.ell_col_d[(long)e * nr + r];
      ev[e] = evm[(long)e * nr + r];
    } else {
      ecol[e] = -1;
      ev[e] = mk<R>((R)0, (R)0);
    }
  }
  int pidx[PW], pcol[PW];
  const int p0 = op.pat_ptr_d[s], np = op.pat_ptr_d[s + 1] - p0;
#pragma unroll
  for (int p = 0; p < PW; ++p) {
    pidx[p] = p < np ? p0 + p : -1;
    pcol[p] = p < np ? op.pat_col_d[p0 + p] : 0;
  }
  const cx<R> k0 = ccast<R>(ld2(op.k0_d, k));
  const int G = (bt.B + (1 << bt.lane_shift) - 1) >> bt.lane_shift;
  for (int g = x + 8 * q; g < G; g += 8 * Q) {
    const int b = (g << bt.lane_shift) + lane;
    if (b < bt.B)
      stage_row_reg<R, WR, PW, NT, V, false, HY>(op, bt, stage, yin, yout, b, r, k, k0, ecol, ev, pidx, pcol);
  }
}

// LDS-staged operator-stationary form.  As k_stage_shared, but a tile is
// aux_per_tile whole aux blocks, and for each trajectory group the tile's own
// values and the group's stage coefficients are staged in LDS; the
// system-block pattern values and coefficients of every row are then LDS
// reads instead of L1/L2 gathers.  The LDS is double-buffered and the next
// group's own values and coefficients are loaded into registers while the
// current group computes, so each group costs one barrier and no exposed
// load.  Threads beyond the tile's rows only help stage the coefficients.
// Every thread reaches the barrier of every group iteration.
template <typename R, int V>
struct LdsStage {
  cx<R> yr[V];  // own values of the group (active rows)
  cx<R> c[V];   // one staged coefficient vector of the group (if cidx >= 0)
  int cidx;     // its LDS offset, or -1
};
// The complex64 forms with ELL and pattern widths 4 and up to 2 lanes per
// thread fit 128 VGPRs without spilling: they ask for 4 waves per SIMD (two
// workgroups per CU); the others take what their registers allow.
// (the forms that also add a dense system block need four more registers per
// lane pair: they take what their registers allow rather than spill)
template <typename R, int WR, int PW, int V, bool HY>
constexpr int lds_min_waves() {
  return (!HY && sizeof(R) == 4 && WR == 4 && PW == 4 && V <= 2) ? 4 : 1;
}
template <typename R, int WR, int PW, int NT, int V, bool HY = false>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(lds_min_waves<R, WR, PW, V, HY>()))) void k_stage_lds(const hops_op_t op, const hops_batch_t bt, int stage,
                                                   const cx<R> *__restrict__ yin, cx<R> *__restrict__ yout,
                                                   int n_tiles, int Q, int aux_per_tile) {
  extern __shared__ double lds_raw[];
  const int x = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int q = slot / n_tiles;
  const int tile = slot - q * n_tiles;
  const int n = op.n_state, np = op.n_pat;
  const int ls = bt.lane_shift, T = 1 << ls;
  const int tps = ls - (V == 4 ? 2 : (V == 2 ? 1 : 0));  // log2 threads per row
  const int TR = 1 << tps;
  const int lr = threadIdx.x >> tps;                       // row within the tile
  const int lane = (threadIdx.x & (TR - 1)) * V;
  const int k_first = tile * aux_per_tile;
  const int n_aux = min(aux_per_tile, op.n_hier - k_first);
  const bool active = lr < n_aux * n;
  const int kl = active ? lr / n : 0;
  const int k = k_first + kl, s = active ? lr - kl * n : 0;
  const long nr = (long)n * op.n_hier;
  const long r = (long)k * n + s;
  const int nsets = k_first == 0 ? 2 : 1;  // the k = 0 coefficient set follows the k > 0 set
  // LDS buffer h (0, 1): rows [aux_per_tile n][T], then coefficients [2 np][T]
  const size_t buf = (size_t)T * ((size_t)aux_per_tile * n + 2 * (size_t)np);
  cx<R> *lds = reinterpret_cast<cx<R> *>(lds_raw);
  int ecol[WR];
  cx<R> ev[WR];
  const cx<R> *evm = reinterpret_cast<const cx<R> *>(op.ell_val_d);
#pragma unroll
  for (int e = 0; e < WR; ++e) {
    if (active && e < op.ell_w) {
      ecol[e] = op.ell_col_d[(long)e * nr + r];
      ev[e] = evm[(long)e * nr + r];
    } else {
      ecol[e] = -1;
      ev[e] = mk<R>((R)0, (R)0);
    }
  }
  int pidx[PW], pcol[PW];
  const int p0 = op.pat_ptr_d[s], npr = active ? op.pat_ptr_d[s + 1] - p0 : 0;
#pragma unroll
  for (int p = 0; p < PW; ++p) {
    pidx[p] = p < npr ? p0 + p : -1;
    pcol[p] = p < npr ? op.pat_col_d[p0 + p] : 0;
  }
  const cx<R> k0 = ccast<R>(ld2(op.k0_d, k));
  const cx<R> *cT = reinterpret_cast<const cx<R> *>(bt.coefT_d);
  const uint32_t cs = (uint32_t)bt.coefT_stride;
  const uint32_t nrT = (uint32_t)nr * (uint32_t)T;
  constexpr int NTO = NT >= 3 ? 1 : 0;
  const int G = (bt.B + T - 1) >> ls;
  const int ncv = nsets * np * TR;  // coefficient vectors of a group
  // global loads of group g into registers (coefficient vectors
  // threadIdx.x + 512 j: at most one per thread when ncv <= 512, else a loop)
  auto fetch = [&](int g, LdsStage<R, V> &f) {
    f.cidx = -1;
    if (threadIdx.x < ncv) {
      const int pr = threadIdx.x >> tps, sl = threadIdx.x & (TR - 1);
      const int bb = (g << ls) + sl * V;
      if (bb < bt.B) {
        ldv<0, V>(el(cT, (uint32_t)bb + (uint32_t)pr * cs), f.c);
        f.cidx = pr * T + sl * V;
      }
    }
    const int b = (g << ls) + lane;
    if (active && b < bt.B) ldv<NTO, V>(el(yin, (uint32_t)g * nrT + (uint32_t)r * (uint32_t)T + (uint32_t)lane), f.yr);
  };
  int h = 0;
  LdsStage<R, V> f;
  int g = x + 8 * q;
  if (g < G) fetch(g, f);
  for (; g < G; g += 8 * Q, h ^= 1) {
    cx<R> *ys = lds + h * buf;
    cx<R> *cl = ys + (size_t)aux_per_tile * n * T;
    const int b = (g << ls) + lane;
    const bool run = active && b < bt.B;
    if (run) {
#pragma unroll
      for (int v = 0; v < V; ++v) ys[lr * T + lane + v] = f.yr[v];
    }
    if (f.cidx >= 0) {
#pragma unroll
      for (int v = 0; v < V; ++v) cl[f.cidx + v] = f.c[v];
    }
    for (int i = threadIdx.x + 512; i < ncv; i += 512) {  // groups with more than 512 coefficient vectors
      const int pr = i >> tps, sl = i & (TR - 1);
      const int bb = (g << ls) + sl * V;
      if (bb < bt.B) {
        cx<R> c[V];
        ldv<0, V>(el(cT, (uint32_t)bb + (uint32_t)pr * cs), c);
#pragma unroll
        for (int v = 0; v < V; ++v) cl[pr * T + sl * V + v] = c[v];
      }
    }
    cx<R> yr[V];
#pragma unroll
    for (int v = 0; v < V; ++v) yr[v] = f.yr[v];
    __syncthreads();
    if (g + 8 * Q < G) fetch(g + 8 * Q, f);  // next group, in flight during this one
    if (run)
      stage_row_reg<R, WR, PW, NT, V, true, HY>(op, bt, stage, yin, yout, b, r, k, k0, ecol, ev, pidx, pcol,
                                            ys + (size_t)(kl * n) * T + lane,
                                            cl + (size_t)(k == 0 ? np : 0) * T + lane, yr);
  }
}

// dst[dst_idx[i]] = scale * src[src_idx[i]]: moves the stable (auxiliary,
// state) entries of Phi into a new basis (basis/hops_basis.py:214-274).
__global__ __launch_bounds__(256) void k_permute(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                 const int32_t *__restrict__ si, const int32_t *__restrict__ di,
                                                 int n, double scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 v = src[si[i]];
  dst[di[i]] = make_double2(scale * v.x, scale * v.y);
}

// phi_out[b, k, s] = sum_s' op[s, s'] phi_in[b, k, s'] (HopsTrajectory._operator,
// trajectory/hops_trajectory.py:560-562); one thread per output element.
__global__ __launch_bounds__(256) void k_state_op(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                  const double2 *__restrict__ op, int n, int nh, int B,
                                                  int lane_shift) {
  const long nr = (long)n * nh;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr * B) return;
  const int b = (int)(i / nr);
  const long r = i - (long)b * nr;
  const int k = (int)(r / n), s = (int)(r - (long)k * n);
  const int T = 1 << lane_shift;
  const long base = (long)(b >> lane_shift) * nr * T + (b & (T - 1));
  const double2 *row = in + base + (long)k * n * T;
  double re = 0.0, im = 0.0;
  for (int q = 0; q < n; ++q) {
    const double2 a = op[(long)s * n + q], v = row[(long)q * T];
    re += a.x * v.x - a.y * v.y;
    im += a.x * v.y + a.y * v.x;
  }
  out[base + r * T] = make_double2(re, im);
}

// y0 = complex64(Phi) at the start of a mixed-precision propagate call.
__global__ __launch_bounds__(256) void k_prime_y0(const double2 *__restrict__ phi,
                                                  float2 *__restrict__ y0, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 p = phi[i];
  y0[i] = make_float2((float)p.x, (float)p.y);
}

// End of a propagate call: sigma of the last step's state and psi storage.
__global__ __launch_bounds__(256) void k_final_sigma(const hops_op_t op, const hops_batch_t bt,
                                                     int store_slot) {
  __shared__ double red[16];
  const int n = op.n_state, b = blockIdx.x, tid = threadIdx.x;
  const Lane ln = lane_of(bt, (long)n * op.n_hier, b);
  const double2 *phi = reinterpret_cast<const double2 *>(bt.phi_d) + ln.base;
  const int T = ln.T;
  zd acc = mk<double>(0.0, 0.0);
  for (int s = tid; s < n; s += blockDim.x) {
    const double2 v = phi[(long)s * T];
    acc.x += v.x * v.x + v.y * v.y;
  }
  acc = block_sum(acc, red);
  const double sigma = bt.eom == HOPS_EOM_NORMALIZED_NONLINEAR ? 1.0 / sqrt(acc.x) : 1.0;
  if (store_slot >= 0 && bt.psi_store_d) {
    double *dst = bt.psi_store_d + 2 * ((long)b * bt.store_stride + store_slot) * n;
    for (int s = tid; s < n; s += blockDim.x) {
      const double2 v = phi[(long)s * T];
      dst[2 * s] = sigma * v.x;
      dst[2 * s + 1] = sigma * v.y;
    }
  }
  if (tid == 0) bt.ws_d[(long)b * bt.ws_stride + 4 * op.n_mode_abs] = sigma;
}

__global__ __launch_bounds__(256) void k_final_scale(const hops_op_t op, const hops_batch_t bt) {
  const int t = bt.tile_d[blockIdx.x];
  if (t < 0) return;
  const int g = t >> 16, tile = t & 0xffff;
  const int lane = threadIdx.x & ((1 << bt.lane_shift) - 1);
  const int b = (g << bt.lane_shift) + lane;
  if (b >= bt.B) return;
  const long nr = (long)op.n_state * op.n_hier;
  const long r = (long)tile * (bt.tile_rows >> bt.lane_shift) + (threadIdx.x >> bt.lane_shift);
  if (r >= nr) return;
  const double sigma = bt.ws_d[(long)b * bt.ws_stride + 4 * op.n_mode_abs];
  const Lane ln = lane_of(bt, nr, b);
  double2 *phi = reinterpret_cast<double2 *>(bt.phi_d);
  double2 p = phi[ln.base + r * ln.T];
  p.x *= sigma;
  p.y *= sigma;
  phi[ln.base + r * ln.T] = p;
}

// Optional per-launch timing: an event is recorded on the launch stream after
// every kernel; the gap between consecutive events is that kernel's time.
struct Profile {
  bool on = false;
  std::vector<hipEvent_t> ev;
  std::vector<int> kind;  // 0 prologue, 1 stage, 2 other, 3 dense system block
  double ms[4] = {0, 0, 0, 0};
  long long n[4] = {0, 0, 0, 0};
};
// per host thread: concurrent callers on other threads keep their own events
// and counters
thread_local Profile g_prof;
bool g_stationary = true;
int g_chunk_traj = 0;
int g_nt_stream = 0;  // nontemporal hint level of k_stage_shared (see stage_row_reg)
int g_groups_q = 0;        // trajectory-group split of k_stage_shared (0: automatic)
int g_vec_lanes = 1;       // trajectory lanes per thread of k_stage_shared (1, 2 or 4)
int g_lds_stage = 0;       // 1: LDS-staged form k_stage_lds where it fits
int g_fuse_linear = 1;     // 1: one prologue launch per LINEAR step (k_prologue_linear4)
int g_graph_steps = 16;    // RK4 steps per captured HIP graph (0: every kernel launched eagerly)
int g_resident_rows = 1024;  // rows per trajectory up to which k_steps_resident runs a whole call

__global__ __launch_bounds__(256) void k_zero_words(unsigned long long *p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0ull;
}

// counter {noise index offset, store slot offset} read by the prologues of a
// replayed graph; the graph's last node advances it by the captured steps
__global__ void k_step_advance(int *ctr, int d_tidx, int d_slot) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctr[0] += d_tidx;
    ctr[1] += d_slot;
  }
}

// Capture happens on a non-blocking stream created for the call (the
// caller's stream may be the legacy default stream, which cannot capture,
// and concurrent callers on other host threads capture on streams of their
// own); the graph is then launched on the caller's stream.  An executable
// graph is destroyed only after the event recorded behind its last launch
// has completed.
struct RetiredGraph {
  hipGraphExec_t ex;
  hipEvent_t done;
};
std::vector<RetiredGraph> g_retired;
std::mutex g_graph_mu;  // the retired list and per-device scratch addresses, across host threads
constexpr int kMaxDevices = 64;

struct CaptureStream {
  hipStream_t s = nullptr;
  ~CaptureStream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

void reap_graphs() {
  std::lock_guard<std::mutex> lk(g_graph_mu);
  size_t w = 0;
  for (size_t i = 0; i < g_retired.size(); ++i) {
    if (hipEventQuery(g_retired[i].done) == hipSuccess) {
      (void)hipGraphExecDestroy(g_retired[i].ex);
      (void)hipEventDestroy(g_retired[i].done);
    } else {
      g_retired[w++] = g_retired[i];
    }
  }
  g_retired.resize(w);
}

int retire_graph(hipGraphExec_t ex, hipStream_t st) {
  hipEvent_t e = nullptr;
  hipError_t rc = hipEventCreateWithFlags(&e, hipEventDisableTiming);
  if (rc == hipSuccess) rc = hipEventRecord(e, st);
  if (rc != hipSuccess) {
    // no completion event: wait for the stream before the graph goes
    (void)hipStreamSynchronize(st);
    (void)hipGraphExecDestroy(ex);
    if (e) (void)hipEventDestroy(e);
    return fail("graph retirement event: %s", hipGetErrorString(rc));
  }
  std::lock_guard<std::mutex> lk(g_graph_mu);
  g_retired.push_back({ex, e});
  return 0;
}

// Scratch of the allocation-free entry point hops_propagate_noalloc: the
// coefficient tables of the fused LINEAR prologue.  One copy per device, held
// until the process ends; host code uses only the per-device address taken
// with hipGetSymbolAddress, never the array's name.
constexpr size_t kNoAllocScratchBytes = HOPS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
void *g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch(void **out) {
  std::lock_guard<std::mutex> lk(g_graph_mu);
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail("device %d outside [0, %d)", dev, kMaxDevices);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail("hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int prof_mark(bool on, hipStream_t st, int kind, size_t *used) {
  if (!on) return 0;
  if (*used >= g_prof.ev.size()) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    g_prof.ev.push_back(e);
    g_prof.kind.push_back(0);
  }
  g_prof.kind[*used] = kind;
  HIP_TRY(hipEventRecord(g_prof.ev[*used], st));
  ++*used;
  return 0;
}

int prof_collect(size_t used) {
  if (!g_prof.on || used < 2) return 0;
  HIP_TRY(hipEventSynchronize(g_prof.ev[used - 1]));
  for (size_t i = 1; i < used; ++i) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, g_prof.ev[i - 1], g_prof.ev[i]));
    g_prof.ms[g_prof.kind[i]] += ms;
    g_prof.n[g_prof.kind[i]] += 1;
  }
  return 0;
}

int check_op(const hops_op_t *op) {
  if (!op) return fail("op is NULL");
  if (op->n_state <= 0 || op->n_hier <= 0) return fail("empty basis (n_state=%d n_hier=%d)", op->n_state, op->n_hier);
  if ((long)op->n_state * op->n_hier >= (1L << 31)) return fail("basis too large for int32 rows");
  if (op->ell_w < 0 || op->n_pat < 0 || op->n_act < 0) return fail("negative sizes");
  if (op->hd_d && (op->hd_n < op->n_state || op->hd_n % 16 != 0))
    return fail("dense system block of side %d for %d states (a multiple of 16 >= n_state)", op->hd_n, op->n_state);
  return 0;
}

// ext_scr: caller-provided scratch of kNoAllocScratchBytes (the
// allocation-free form: no stream-ordered allocation, no graph capture, no
// profiler events); nullptr: scratch from hipMallocAsync
template <typename R>
int propagate_t(const hops_op_t *op, const hops_batch_t *bt, int it0, int n_steps, int store0,
                hipStream_t st, void *ext_scr) {
  const bool noalloc = ext_scr != nullptr;
  const bool prof = g_prof.on && !noalloc;
  const int n = op->n_state;
  const long nr = (long)n * op->n_hier;
  const size_t shm = sizeof(double) * (2 * (size_t)n + 8 * (size_t)op->n_act + 2 * (size_t)op->n_mode_abs + 16);
  if (shm > 160 * 1024) return fail("prologue LDS %zu B exceeds 160 KiB", shm);
  if (shm > 64 * 1024) {
    HIP_TRY(hipFuncSetAttribute((const void *)k_prologue<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    HIP_TRY(hipFuncSetAttribute((const void *)k_prologue_linear4<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shm));
    HIP_TRY(hipFuncSetAttribute((const void *)k_steps_resident<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shm));
  }
  const int ns = bt->noise_step;
  if (it0 < 0 || it0 + (long)n_steps * 2 * ns >= bt->n_t)
    return fail("noise index range [%d, %ld] outside [0, %d)", it0, it0 + (long)n_steps * 2 * ns, bt->n_t);
  const bool dense = op->hd_d != nullptr;
  if (dense != (bt->hy_d != nullptr)) return fail("dense system block needs its HY pool (and only it)");
  // scratch of the fused LINEAR prologue: the coefficient tables of stages 1..3
  const bool fuse = g_fuse_linear && bt->eom == HOPS_EOM_LINEAR;
  const size_t coefT_blk = sizeof(cx<R>) * (size_t)bt->coefT_stride *
                           (2 * (size_t)op->n_pat + (size_t)op->n_mode + 1);
  const size_t coef_blk = (size_t)bt->B * (size_t)bt->coef_stride;
  const size_t fuse_bytes = fuse && n_steps > 0 ? 3 * coefT_blk + 3 * coef_blk * sizeof(double) : 0;
  if (noalloc && fuse_bytes > kNoAllocScratchBytes)
    return fail("hops_propagate_noalloc needs %zu B of scratch for this batch, more than its %zu B", fuse_bytes,
                kNoAllocScratchBytes);
  const int G = (bt->B + (1 << bt->lane_shift) - 1) >> bt->lane_shift;
  cx<R> *Y0 = reinterpret_cast<cx<R> *>(bt->precision == HOPS_PREC_MIXED ? bt->y0_d : bt->phi_d);
  cx<R> *ya = reinterpret_cast<cx<R> *>(bt->ya_d);
  cx<R> *yb = reinterpret_cast<cx<R> *>(bt->yb_d);
  if (bt->precision == HOPS_PREC_MIXED) {
    const long tot = nr * ((long)G << bt->lane_shift);
    k_prime_y0<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(reinterpret_cast<const double2 *>(bt->phi_d),
                                                            reinterpret_cast<float2 *>(bt->y0_d), tot);
    HIP_TRY(hipGetLastError());
  }
  const int toff[4] = {0, ns, ns, 2 * ns};
  if (n_steps > 0 && nr <= g_resident_rows && !dense) {
    size_t used = 0;
    if (int rc = prof_mark(prof, st, 2, &used)) return rc;
    k_steps_resident<R><<<bt->B, 256, shm, st>>>(*op, *bt, it0, n_steps, store0, Y0, ya, yb);
    HIP_TRY(hipGetLastError());
    if (int rc = prof_mark(prof, st, 1, &used)) return rc;
    if (int rc = prof_collect(used)) return rc;
    const int slot = bt->psi_store_d ? store0 + n_steps - 1 : -1;
    k_final_sigma<<<bt->B, 256, 0, st>>>(*op, *bt, slot);
    HIP_TRY(hipGetLastError());
    if (bt->eom == HOPS_EOM_NORMALIZED_NONLINEAR) {
      k_final_scale<<<bt->n_tiles, 256, 0, st>>>(*op, *bt);
      HIP_TRY(hipGetLastError());
    }
    return 0;
  }
  // operator-stationary stage kernel when every XCD has a trajectory group and
  // the row's ELL and pattern entries fit in registers
  const int max_pat = op->pat_w;
  const int T = 1 << bt->lane_shift;
  int chunk = g_chunk_traj;
  if (chunk > 0) chunk = (chunk / (8 * T)) * 8 * T;  // whole groups, one per XCD at least
  if (chunk <= 0 || chunk >= bt->B) chunk = bt->B;
  const int Gc = (chunk + T - 1) / T;
  int sh_w = 0, sh_p = 0, sh_tiles = 0, sh_q = 1;
  unsigned sh_grid = 0;
  int lds_apt = 0, lds_tiles = 0;
  size_t lds_bytes = 0;
  unsigned lds_grid = 0;
  // lanes per thread of the stationary kernel: at most the interleave, and a
  // divisor of every (sub-)batch size so no vector straddles the last lane
  int vl = g_vec_lanes;
  while (vl > 1 && (vl > T || bt->B % vl != 0 || chunk % vl != 0)) vl >>= 1;
  if (dense && vl > 2) vl = 2;  // the forms with a dense block take at most 2 (the grid below follows vl)
  // the stationary kernel addresses its pools with 32-bit byte offsets
  const bool pools32 = (double)nr * (double)((long)G << bt->lane_shift) * 16.0 < 4294967296.0;
  // (a dense system block: the stationary forms add its (-iH) Y, built for
  // pattern width 4, which the L terms alone fill at most)
  if (g_stationary && pools32 && Gc >= 8 && op->ell_w <= 16 && max_pat > 0 && max_pat <= (dense ? 4 : 8)) {
    sh_w = op->ell_w <= 4 ? 4 : (op->ell_w <= 8 ? 8 : 16);
    sh_p = max_pat <= 4 ? 4 : 8;
    const long rows_per_tile = (512L * vl) >> bt->lane_shift;
    sh_tiles = (int)((nr + rows_per_tile - 1) / rows_per_tile);
    const int per_xcd = (Gc + 7) / 8;
    sh_q = g_groups_q > 0 ? std::min(g_groups_q, per_xcd) : (per_xcd + 31) / 32;
    sh_grid = (unsigned)(8L * sh_q * sh_tiles);
    // LDS-staged form: whole aux blocks per tile, staged rows and
    // coefficients within 64 KiB
    const int apt = (int)(rows_per_tile / n);
    if (g_lds_stage && apt >= 1) {
      const size_t lds = 2 * sizeof(cx<R>) * (size_t)T * ((size_t)apt * n + 2 * (size_t)op->n_pat);
      if (lds <= 64 * 1024) {
        lds_apt = apt;
        lds_bytes = lds;
        lds_tiles = (op->n_hier + apt - 1) / apt;
        lds_grid = (unsigned)(8L * sh_q * lds_tiles);
      }
    }
  } else {
    chunk = bt->B;  // chunks need the stationary kernel (no tile list per chunk)
  }
  if (dense) chunk = bt->B;  // the dense block's launch covers the whole batch
  // fused LINEAR prologue: stream-ordered scratch for the coefficient
  // tables of stages 1..3, freed on the same stream when the call ends (or
  // the caller's scratch of the allocation-free form)
  // graph replay: step 0 eagerly, then steps 1 .. in launches of K captured
  // steps whose prologues add the device step counter (zeroed below)
  const int K = g_graph_steps;
  // (a call replays its graph at least 4 times, or the capture and
  // instantiation cost more than the launches they save)
  bool graph = !noalloc && K > 0 && !prof && chunk >= bt->B && n_steps - 1 >= 4 * K;
  if (graph) {
    // a caller that captures this call into a graph of its own gets eager
    // launches (no capture inside a capture)
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    graph = hipStreamIsCapturing(st, &cap) == hipSuccess && cap == hipStreamCaptureStatusNone;
  }
  const size_t ctr_off = (fuse_bytes + 255) & ~(size_t)255;
  const size_t scr_bytes = graph ? ctr_off + 256 : fuse_bytes;
  struct StreamScratch {
    void *p = nullptr;
    hipStream_t s;
    ~StreamScratch() {
      if (p) (void)hipFreeAsync(p, s);
    }
  } scr;
  scr.s = st;
  void *scr_base = nullptr;  // no graph in the allocation-free form: scr_bytes == fuse_bytes there
  if (scr_bytes > 0) {
    if (noalloc) {
      scr_base = ext_scr;
    } else {
      HIP_TRY(hipMallocAsync(&scr.p, scr_bytes, st));
      scr_base = scr.p;
    }
    // cleared by a kernel, not a memset node, should the caller capture
    const long words = (long)(scr_bytes / 8);
    k_zero_words<<<(unsigned)std::min<long>((words + 255) / 256, 4096), 256, 0, st>>>(
        reinterpret_cast<unsigned long long *>(scr_base), words);
    HIP_TRY(hipGetLastError());
  }
  char *coefT_s = reinterpret_cast<char *>(scr_base);
  double *coef_s = reinterpret_cast<double *>(coefT_s + 3 * coefT_blk);
  int *step_ctr = graph ? reinterpret_cast<int *>(coefT_s + ctr_off) : nullptr;
  size_t used = 0;
  if (int rc = prof_mark(prof, st, 2, &used)) return rc;
  // one RK4 step of a (sub-)batch on stream ls; slot: psi store slot written
  // at stage 0; sc: step counter of a captured graph (nullptr: eager)
  auto one_step = [&](const hops_batch_t &sb, int step, hipStream_t ls, const int *sc) -> int {
    const int itb = it0 + step * 2 * ns;
    // sub-batch offsets into the (scratch) tables
    const size_t offT = (size_t)((const char *)sb.coefT_d - (const char *)bt->coefT_d);
    const size_t offC = (size_t)(sb.coef_d - bt->coef_d);
    if (fuse) {
      const cx<R> *yin0 = reinterpret_cast<const cx<R> *>(sb.precision == HOPS_PREC_MIXED ? sb.y0_d : sb.phi_d);
      const int slot = (step > 0 && sb.psi_store_d) ? store0 + step - 1 : -1;
      k_prologue_linear4<R><<<dim3(sb.B, 4), 256, shm, ls>>>(*op, sb, itb, ns, yin0, slot, coefT_s + offT, coefT_blk,
                                                              coef_s + offC, coef_blk, sc);
      HIP_TRY(hipGetLastError());
      if (int rc = prof_mark(prof, ls, 0, &used)) return rc;
    }
    for (int i = 0; i < 4; ++i) {
      const cx<R> *yin = i == 0 ? reinterpret_cast<const cx<R> *>(sb.precision == HOPS_PREC_MIXED ? sb.y0_d : sb.phi_d)
                                : reinterpret_cast<const cx<R> *>(i == 2 ? sb.yb_d : sb.ya_d);
      cx<R> *yout = reinterpret_cast<cx<R> *>(i == 1 ? sb.yb_d : sb.ya_d);
      const int slot = (i == 0 && step > 0 && sb.psi_store_d) ? store0 + step - 1 : -1;
      hops_batch_t sbi = sb;
      if (fuse) {
        if (i > 0) sbi.coefT_d = coefT_s + (size_t)(i - 1) * coefT_blk + offT;
      } else {
        k_prologue<R><<<sb.B, 256, shm, ls>>>(*op, sb, i, itb + toff[i], yin, slot, sc);
        HIP_TRY(hipGetLastError());
        if (int rc = prof_mark(prof, ls, 0, &used)) return rc;
      }
#define HOPS_SHV(W, P, NT)                                                                       \
  do {                                                                                           \
    if (lds_apt > 0) {                                                                           \
      if (vl == 4) k_stage_lds<R, W, P, NT, 4><<<lds_grid, 512, lds_bytes, ls>>>(*op, sbi, i, yin, yout, lds_tiles, sh_q, lds_apt); \
      else if (vl == 2) k_stage_lds<R, W, P, NT, 2><<<lds_grid, 512, lds_bytes, ls>>>(*op, sbi, i, yin, yout, lds_tiles, sh_q, lds_apt); \
      else k_stage_lds<R, W, P, NT, 1><<<lds_grid, 512, lds_bytes, ls>>>(*op, sbi, i, yin, yout, lds_tiles, sh_q, lds_apt); \
    } else if (vl == 4) k_stage_shared<R, W, P, NT, 4><<<sh_grid, 512, 0, ls>>>(*op, sbi, i, yin, yout, sh_tiles, sh_q); \
    else if (vl == 2) k_stage_shared<R, W, P, NT, 2><<<sh_grid, 512, 0, ls>>>(*op, sbi, i, yin, yout, sh_tiles, sh_q); \
    else k_stage_shared<R, W, P, NT, 1><<<sh_grid, 512, 0, ls>>>(*op, sbi, i, yin, yout, sh_tiles, sh_q); \
  } while (0)
#define HOPS_SH(W, P)                                                                            \
  do {                                                                                           \
    if (g_nt_stream == 1) HOPS_SHV(W, P, 1);                                                     \
    else if (g_nt_stream == 2) HOPS_SHV(W, P, 2);                                                \
    else if (g_nt_stream == 3) HOPS_SHV(W, P, 3);                                                \
    else if (g_nt_stream == 4) HOPS_SHV(W, P, 4);                                                \
    else HOPS_SHV(W, P, 0);                                                                      \
  } while (0)
// with the dense block: hint levels 0 and 2 (a level only changes what L2
// keeps), 1 or 2 lanes per thread (vl was capped at 2 before the grid)
#define HOPS_SHD(W, NT)                                                                          \
  do {                                                                                           \
    if (lds_apt > 0) {                                                                           \
      if (vl >= 2) k_stage_lds<R, W, 4, NT, 2, true><<<lds_grid, 512, lds_bytes, ls>>>(*op, sbi, i, yin, yout, lds_tiles, sh_q, lds_apt); \
      else k_stage_lds<R, W, 4, NT, 1, true><<<lds_grid, 512, lds_bytes, ls>>>(*op, sbi, i, yin, yout, lds_tiles, sh_q, lds_apt); \
    } else if (vl >= 2) k_stage_shared<R, W, 4, NT, 2, true><<<sh_grid, 512, 0, ls>>>(*op, sbi, i, yin, yout, sh_tiles, sh_q); \
    else k_stage_shared<R, W, 4, NT, 1, true><<<sh_grid, 512, 0, ls>>>(*op, sbi, i, yin, yout, sh_tiles, sh_q); \
  } while (0)
#define HOPS_SHD_NT(W)                                                                           \
  do {                                                                                           \
    if (g_nt_stream >= 2) HOPS_SHD(W, 2);                                                        \
    else HOPS_SHD(W, 0);                                                                         \
  } while (0)
      if (dense) {
        if (int rc = launch_dense_h<R>(*op, sbi, G, yin, reinterpret_cast<cx<R> *>(sbi.hy_d), ls)) return rc;
        if (int rc = prof_mark(prof, ls, 3, &used)) return rc;
      }
      if (dense && sh_w == 4 && sh_p == 4) HOPS_SHD_NT(4);
      else if (dense && sh_w == 8 && sh_p == 4) HOPS_SHD_NT(8);
      else if (dense && sh_w == 16 && sh_p == 4) HOPS_SHD_NT(16);
      else if (dense) k_stage<R><<<sb.n_tiles, 256, 0, ls>>>(*op, sbi, i, yin, yout);
      else if (sh_w == 4 && sh_p == 4) HOPS_SH(4, 4);
      else if (sh_w == 4 && sh_p == 8) HOPS_SH(4, 8);
      else if (sh_w == 8 && sh_p == 4) HOPS_SH(8, 4);
      else if (sh_w == 8 && sh_p == 8) HOPS_SH(8, 8);
      else if (sh_w == 16 && sh_p == 4) HOPS_SH(16, 4);
      else if (sh_w == 16 && sh_p == 8) HOPS_SH(16, 8);
      else k_stage<R><<<sb.n_tiles, 256, 0, ls>>>(*op, sbi, i, yin, yout);
#undef HOPS_SH
#undef HOPS_SHV
#undef HOPS_SHD
#undef HOPS_SHD_NT
      HIP_TRY(hipGetLastError());
      if (int rc = prof_mark(prof, ls, 1, &used)) return rc;
    }
    return 0;
  };
  // sub-batch of trajectories [b0, b0 + nb): every pointer offset to it
  auto sub_batch = [&](int b0, int nb) {
    hops_batch_t sb = *bt;
    const size_t es = sizeof(cx<R>), pe = (size_t)b0 * nr;  // b0 is a multiple of T
    sb.B = nb;
    sb.phi_d = (char *)bt->phi_d + pe * 16;
    sb.y0_d = bt->precision == HOPS_PREC_MIXED ? (char *)bt->y0_d + pe * 8 : bt->y0_d;
    sb.ya_d = (char *)bt->ya_d + pe * es;
    sb.yb_d = (char *)bt->yb_d + pe * es;
    sb.acc_d = (char *)bt->acc_d + pe * es;
    const size_t zes = bt->noise_c128 ? 16 : 8;
    sb.z1_d = (const char *)bt->z1_d + (size_t)b0 * op->n_act * bt->n_t * zes;
    if (bt->z2_d) sb.z2_d = (const char *)bt->z2_d + (size_t)b0 * op->n_act * bt->n_t * zes;
    sb.zmem_d = bt->zmem_d + (size_t)b0 * 2 * op->n_mode_abs;
    sb.ws_d = bt->ws_d + (size_t)b0 * bt->ws_stride;
    sb.coef_d = bt->coef_d + (size_t)b0 * bt->coef_stride;
    sb.coefT_d = (char *)bt->coefT_d + (size_t)b0 * es * 1;
    sb.sigT_d = bt->sigT_d + b0;
    if (bt->psi_store_d) sb.psi_store_d = bt->psi_store_d + (size_t)b0 * bt->store_stride * 2 * op->n_state;
    return sb;
  };
  if (graph) {
    if (int rc = one_step(*bt, 0, st, nullptr)) return rc;
    reap_graphs();
    CaptureStream cap;
    HIP_TRY(hipStreamCreateWithFlags(&cap.s, hipStreamNonBlocking));
    hipStream_t cs = cap.s;
    HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int j = 0; j < K && !rc; ++j) rc = one_step(*bt, 1 + j, cs, step_ctr);
    if (!rc) {
      k_step_advance<<<1, 64, 0, cs>>>(step_ctr, K * 2 * ns, K);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) rc = fail("k_step_advance launch: %s", hipGetErrorString(e));
    }
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(cs, &g);
    if (rc || ce != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : fail("hipStreamEndCapture failed: %s", hipGetErrorString(ce));
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiateWithFlags(&ex, g, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) return fail("hipGraphInstantiate failed: %s", hipGetErrorString(ie));
    const int n_rep = (n_steps - 1) / K;
    hipError_t le = hipSuccess;
    for (int c = 0; c < n_rep && le == hipSuccess; ++c) le = hipGraphLaunch(ex, st);
    if (int rr = retire_graph(ex, st)) return rr;
    if (le != hipSuccess) return fail("hipGraphLaunch failed: %s", hipGetErrorString(le));
    for (int step = 1 + n_rep * K; step < n_steps; ++step)
      if (int r2 = one_step(*bt, step, st, nullptr)) return r2;
  } else if (chunk >= bt->B) {
    for (int step = 0; step < n_steps; ++step)
      if (int rc = one_step(*bt, step, st, nullptr)) return rc;
  } else {
    // all steps of one chunk before the next: the chunk's pools stay in the
    // last-level cache across the whole propagation
    for (int b0 = 0; b0 < bt->B; b0 += chunk) {
      const hops_batch_t sb = sub_batch(b0, std::min(chunk, bt->B - b0));
      for (int step = 0; step < n_steps; ++step)
        if (int rc = one_step(sb, step, st, nullptr)) return rc;
    }
  }
  if (int rc = prof_collect(used)) return rc;
  const int slot = (n_steps > 0 && bt->psi_store_d) ? store0 + n_steps - 1 : -1;
  k_final_sigma<<<bt->B, 256, 0, st>>>(*op, *bt, slot);
  HIP_TRY(hipGetLastError());
  if (bt->eom == HOPS_EOM_NORMALIZED_NONLINEAR) {
    k_final_scale<<<bt->n_tiles, 256, 0, st>>>(*op, *bt);
    HIP_TRY(hipGetLastError());
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Ragged batch ("slots").  Every slot is one trajectory with its own basis:
// its own hops_op_t and a B = 1, lane_shift = 0 hops_batch_t, both in device
// arrays indexed by slot.  One launch covers every listed slot; a tile is 256
// rows of one slot, encoded slot << kSlotShift | tile.
// ---------------------------------------------------------------------------
constexpr int kSlotShift = 20;
constexpr uint32_t kTileMask = (1u << kSlotShift) - 1;

// stage input / output pools (stage < 0: derivative evaluation ya -> yb)
template <typename R>
__device__ __forceinline__ const cx<R> *slot_in(const hops_batch_t &bt, int stage) {
  if (stage < 0) return reinterpret_cast<const cx<R> *>(bt.ya_d);
  if (stage == 0)
    return reinterpret_cast<const cx<R> *>(bt.precision == HOPS_PREC_MIXED ? bt.y0_d : bt.phi_d);
  return reinterpret_cast<const cx<R> *>(stage == 2 ? bt.yb_d : bt.ya_d);
}
template <typename R>
__device__ __forceinline__ cx<R> *slot_out(const hops_batch_t &bt, int stage) {
  if (stage < 0) return reinterpret_cast<cx<R> *>(bt.yb_d);
  return reinterpret_cast<cx<R> *>(stage == 1 ? bt.yb_d : bt.ya_d);
}

__constant__ int c_toff[4] = {0, 1, 1, 2};

// it[slot]: noise index of the slot's current step start
template <typename R>
__global__ __launch_bounds__(256) void k_prologue_slots(const hops_op_t *__restrict__ ops,
                                                        const hops_batch_t *__restrict__ bts,
                                                        const int32_t *__restrict__ list,
                                                        const int32_t *__restrict__ it, int stage) {
  const int slot = list[blockIdx.x];
  const hops_op_t &op = ops[slot];
  const hops_batch_t &bt = bts[slot];
  const int tidx = it[slot] + (stage < 0 ? 0 : c_toff[stage] * bt.noise_step);
  prologue_body<R>(op, bt, 0, stage, tidx, slot_in<R>(bt, stage), -1);
}

template <typename R>
__global__ __launch_bounds__(256) void k_stage_slots(const hops_op_t *__restrict__ ops,
                                                     const hops_batch_t *__restrict__ bts,
                                                     const uint32_t *__restrict__ tiles, int stage) {
  const uint32_t t = tiles[blockIdx.x];
  const int slot = (int)(t >> kSlotShift);
  const hops_op_t &op = ops[slot];
  const hops_batch_t &bt = bts[slot];
  const int n = op.n_state;
  const long nr = (long)n * op.n_hier;
  const long r = (long)(t & kTileMask) * 256 + threadIdx.x;
  if (r >= nr) return;
  const int k = (int)(r / n), s = (int)(r - (long)k * n);
  stage_row<R, 0>(op, bt, stage, slot_in<R>(bt, stage), slot_out<R>(bt, stage), 0, r, k, s,
                  ccast<R>(ld2(op.k0_d, k)), nullptr, nullptr);
}

// complex64 copy of Phi: the stage-0 input of a mixed-precision step
__global__ __launch_bounds__(256) void k_prime_slots(const hops_op_t *__restrict__ ops,
                                                     const hops_batch_t *__restrict__ bts,
                                                     const uint32_t *__restrict__ tiles) {
  const uint32_t t = tiles[blockIdx.x];
  const int slot = (int)(t >> kSlotShift);
  const hops_op_t &op = ops[slot];
  const hops_batch_t &bt = bts[slot];
  const long nr = (long)op.n_state * op.n_hier;
  const long r = (long)(t & kTileMask) * 256 + threadIdx.x;
  if (r >= nr) return;
  const double2 p = reinterpret_cast<const double2 *>(bt.phi_d)[r];
  reinterpret_cast<float2 *>(bt.y0_d)[r] = make_float2((float)p.x, (float)p.y);
}

// end of a step: Phi <- Phi / |psi| for NORMALIZED NONLINEAR
// (trajectory/hops_trajectory.py:616-634); one workgroup per slot
__global__ __launch_bounds__(256) void k_final_slots(const hops_op_t *__restrict__ ops,
                                                     const hops_batch_t *__restrict__ bts,
                                                     const int32_t *__restrict__ list) {
  __shared__ double red[16];
  const int slot = list[blockIdx.x];
  const hops_op_t &op = ops[slot];
  const hops_batch_t &bt = bts[slot];
  if (bt.eom != HOPS_EOM_NORMALIZED_NONLINEAR) return;  // uniform over the block
  const int n = op.n_state;
  double2 *phi = reinterpret_cast<double2 *>(bt.phi_d);
  zd acc = mk<double>(0.0, 0.0);
  for (int s = threadIdx.x; s < n; s += blockDim.x) acc.x += phi[s].x * phi[s].x + phi[s].y * phi[s].y;
  acc = block_sum(acc, red);
  const double sigma = 1.0 / sqrt(acc.x);
  const long nr = (long)n * op.n_hier;
  for (long i = threadIdx.x; i < nr; i += blockDim.x) {
    double2 p = phi[i];
    p.x *= sigma;
    p.y *= sigma;
    phi[i] = p;
  }
}

// Pn = Phi / |psi|: into ya (stage precision, the derivative input) and in
// FP64 into out + off[slot]; |psi| into nrm[slot] (basis/hops_basis.py:157).
// psi_only: copy psi = Phi[:n_state] to out + off[slot] and nothing else.
template <typename R>
__global__ __launch_bounds__(256) void k_pn_slots(const hops_op_t *__restrict__ ops,
                                                  const hops_batch_t *__restrict__ bts,
                                                  const int32_t *__restrict__ list,
                                                  const int64_t *__restrict__ off, double2 *__restrict__ out,
                                                  double *__restrict__ nrm, int psi_only) {
  __shared__ double red[16];
  const int slot = list[blockIdx.x];
  const hops_op_t &op = ops[slot];
  const hops_batch_t &bt = bts[slot];
  const int n = op.n_state;
  const double2 *phi = reinterpret_cast<const double2 *>(bt.phi_d);
  double2 *o = out + off[slot];
  if (psi_only) {
    for (int s = threadIdx.x; s < n; s += blockDim.x) o[s] = phi[s];
    return;
  }
  zd acc = mk<double>(0.0, 0.0);
  for (int s = threadIdx.x; s < n; s += blockDim.x) acc.x += phi[s].x * phi[s].x + phi[s].y * phi[s].y;
  acc = block_sum(acc, red);
  const double norm = sqrt(acc.x);
  if (threadIdx.x == 0) nrm[slot] = norm;
  cx<R> *ya = reinterpret_cast<cx<R> *>(bt.ya_d);
  const long nr = (long)n * op.n_hier;
  for (long i = threadIdx.x; i < nr; i += blockDim.x) {
    const double2 p = phi[i];
    double2 q;
    q.x = p.x / norm;
    q.y = p.y / norm;
    o[i] = q;
    ya[i] = mk<R>((R)q.x, (R)q.y);
  }
}

// derivative input restricted to the stable auxiliaries: ya[k, s] = Pn[k, s]
// if mask[k] else 0 (basis_functions_adaptive.py:10-74 with list_index_aux_stable)
template <typename R>
__global__ __launch_bounds__(256) void k_mask_slots(const hops_op_t *__restrict__ ops,
                                                    const hops_batch_t *__restrict__ bts,
                                                    const uint32_t *__restrict__ tiles,
                                                    const int64_t *__restrict__ off,
                                                    const double2 *__restrict__ pn,
                                                    const int64_t *__restrict__ moff,
                                                    const uint8_t *__restrict__ mask) {
  const uint32_t t = tiles[blockIdx.x];
  const int slot = (int)(t >> kSlotShift);
  const hops_op_t &op = ops[slot];
  const hops_batch_t &bt = bts[slot];
  const int n = op.n_state;
  const long nr = (long)n * op.n_hier;
  const long r = (long)(t & kTileMask) * 256 + threadIdx.x;
  if (r >= nr) return;
  const int k = (int)(r / n);
  const double2 q = pn[off[slot] + r];
  const bool keep = mask[moff[slot] + k] != 0;
  reinterpret_cast<cx<R> *>(bt.ya_d)[r] = keep ? mk<R>((R)q.x, (R)q.y) : mk<R>((R)0, (R)0);
}

// |D/hbar + P/dt|^2 summed over the states of each auxiliary (mode 0) or over
// the masked auxiliaries of each state (mode 1); D = yb, P = Pn (masked)
// (basis/hops_basis.py:560-665, :795-912, basis_functions_adaptive.py:10-74)
template <typename R>
__global__ __launch_bounds__(256) void k_err_slots(const hops_op_t *__restrict__ ops,
                                                   const hops_batch_t *__restrict__ bts,
                                                   const int32_t *__restrict__ list,
                                                   const int64_t *__restrict__ off,
                                                   const double2 *__restrict__ pn,
                                                   const int64_t *__restrict__ eoff, double *__restrict__ eout,
                                                   const double *__restrict__ dt, double hbar, int mode,
                                                   const int64_t *__restrict__ moff,
                                                   const uint8_t *__restrict__ mask) {
#pragma clang fp contract(off)
  const int slot = list[blockIdx.x];
  const hops_op_t &op = ops[slot];
  const hops_batch_t &bt = bts[slot];
  const int n = op.n_state, nh = op.n_hier;
  const cx<R> *d = reinterpret_cast<const cx<R> *>(bt.yb_d);
  const double2 *p = pn + off[slot];
  const double tau = dt[slot];
  double *e = eout + eoff[slot];
  if (mode == 0) {
    for (int k = threadIdx.x; k < nh; k += blockDim.x) {
      double acc = 0.0;
      for (int s = 0; s < n; ++s) {
        const long r = (long)k * n + s;
        const double vx = (double)d[r].x / hbar + p[r].x / tau;
        const double vy = (double)d[r].y / hbar + p[r].y / tau;
        acc += vx * vx + vy * vy;
      }
      e[k] = acc;
    }
  } else {
    // per state: g = blockDim / n threads split the auxiliaries (k = j, j+g,
    // ...), then one thread adds the g partial sums in a fixed order
    __shared__ double part[256];
    const uint8_t *m = mask + moff[slot];
    for (int s0 = 0; s0 < n; s0 += blockDim.x) {
      const int ns = min(n - s0, (int)blockDim.x);
      const int g = blockDim.x / ns;
      const int s = s0 + threadIdx.x / g, j = threadIdx.x % g;
      double acc = 0.0;
      if ((int)threadIdx.x < g * ns)
        for (int k = j; k < nh; k += g) {
          if (!m[k]) continue;
          const long r = (long)k * n + s;
          const double vx = (double)d[r].x / hbar + p[r].x / tau;
          const double vy = (double)d[r].y / hbar + p[r].y / tau;
          acc += vx * vx + vy * vy;
        }
      part[threadIdx.x] = acc;
      __syncthreads();
      if ((int)threadIdx.x < ns) {
        double t = 0.0;
        for (int q = 0; q < g; ++q) t += part[threadIdx.x * g + q];
        e[s0 + threadIdx.x] = t;
      }
      __syncthreads();
    }
  }
}

// super-operator of listed slots (tiles over their rows / one block per slot)
template <typename R>
__global__ __launch_bounds__(256) void k_build_ell_slots(const hops_op_t *__restrict__ ops,
                                                         const uint32_t *__restrict__ tiles) {
  const uint32_t t = tiles[blockIdx.x];
  build_ell_row<R>(ops[t >> kSlotShift], (long)(t & kTileMask) * 256 + threadIdx.x);
}
__global__ __launch_bounds__(256) void k_build_k0_slots(const hops_op_t *__restrict__ ops,
                                                        const int32_t *__restrict__ list) {
  const hops_op_t &op = ops[list[blockIdx.x]];
  for (int k = threadIdx.x; k < op.n_hier; k += blockDim.x) build_k0_aux(op, k);
}

// Phi of a new basis (hops_basis.py:214-274): new row (j, s') takes
// scale * old[(inv_aux[j], inv_state[s'])] when both exist, else 0; tiles run
// over the new rows of each listed slot.
__global__ __launch_bounds__(256) void k_permute_slots(const hops_slot_perm_t *__restrict__ pd,
                                                       const uint32_t *__restrict__ tiles) {
  const uint32_t t = tiles[blockIdx.x];
  const hops_slot_perm_t &p = pd[t >> kSlotShift];
  const long nr = (long)p.n_new * p.nh_new;
  const long r = (long)(t & kTileMask) * 256 + threadIdx.x;
  if (r >= nr) return;
  const int j = (int)(r / p.n_new), s = (int)(r - (long)j * p.n_new);
  const int i = p.inv_aux_d[j], so = p.inv_state_d[s];
  double2 v = make_double2(0.0, 0.0);
  if (i >= 0 && so >= 0) {
    const double2 q = reinterpret_cast<const double2 *>(p.old_phi_d)[(long)i * p.n_old + so];
    v.x = p.scale * q.x;
    v.y = p.scale * q.y;
  }
  reinterpret_cast<double2 *>(p.new_phi_d)[r] = v;
}

// byte segments (16-byte multiples, 16-byte aligned) from a staging area to
// their destinations; one workgroup per segment
__global__ __launch_bounds__(256) void k_copy_segments(const hops_segment_t *__restrict__ segs) {
  const hops_segment_t sg = segs[blockIdx.x];
  const long n = sg.bytes >> 4;
  const uint4 *src = reinterpret_cast<const uint4 *>(sg.src);
  uint4 *dst = reinterpret_cast<uint4 *>(sg.dst);
  for (long i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_gather_u64(const uint64_t *const *__restrict__ src,
                                                   uint64_t *__restrict__ out, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = *src[i];
}

template <typename R>
int slots_step_t(const hops_op_t *ops, const hops_batch_t *bts, const int32_t *list, int n_list,
                 const uint32_t *tiles, int n_tiles, const int32_t *it, size_t shm, bool mixed,
                 hipStream_t st) {
  if (shm > 64 * 1024)
    HIP_TRY(hipFuncSetAttribute((const void *)k_prologue_slots<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shm));
  const bool prof = g_prof.on;
  size_t used = 0;
  if (int rc = prof_mark(prof, st, 2, &used)) return rc;
  if (mixed) {
    k_prime_slots<<<n_tiles, 256, 0, st>>>(ops, bts, tiles);
    HIP_TRY(hipGetLastError());
    if (int rc = prof_mark(prof, st, 2, &used)) return rc;
  }
  for (int i = 0; i < 4; ++i) {
    k_prologue_slots<R><<<n_list, 256, shm, st>>>(ops, bts, list, it, i);
    HIP_TRY(hipGetLastError());
    if (int rc = prof_mark(prof, st, 0, &used)) return rc;
    k_stage_slots<R><<<n_tiles, 256, 0, st>>>(ops, bts, tiles, i);
    HIP_TRY(hipGetLastError());
    if (int rc = prof_mark(prof, st, 1, &used)) return rc;
  }
  k_final_slots<<<n_list, 256, 0, st>>>(ops, bts, list);
  HIP_TRY(hipGetLastError());
  if (int rc = prof_mark(prof, st, 2, &used)) return rc;
  return prof_collect(used);
}

template <typename R>
int slots_deriv_t(const hops_op_t *ops, const hops_batch_t *bts, const int32_t *list, int n_list,
                  const uint32_t *tiles, int n_tiles, const int32_t *it, size_t shm, hipStream_t st) {
  if (shm > 64 * 1024)
    HIP_TRY(hipFuncSetAttribute((const void *)k_prologue_slots<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shm));
  k_prologue_slots<R><<<n_list, 256, shm, st>>>(ops, bts, list, it, -1);
  HIP_TRY(hipGetLastError());
  k_stage_slots<R><<<n_tiles, 256, 0, st>>>(ops, bts, tiles, -1);
  HIP_TRY(hipGetLastError());
  return 0;
}

int slots_args_ok(const void *ops, const void *bts, const void *list, int n_list, const void *tiles,
                  int n_tiles, size_t shm_doubles) {
  if (!ops || !bts || !list || !tiles) return fail("slots: NULL argument");
  if (n_list <= 0 || n_tiles <= 0) return fail("slots: empty slot or tile list");
  if (shm_doubles * sizeof(double) > 160 * 1024) return fail("slots: prologue LDS exceeds 160 KiB");
  return 0;
}

}  // namespace

extern "C" {

const char *hops_last_error(void) { return g_err.c_str(); }

// Shared by the other translation units of the library (noise_fft.hip).
int hops_internal_set_error(const char *msg) {
  g_err = msg;
  return -1;
}

int hops_version(void) { return 4; }

int hops_profile_enable(int on) {
  g_prof.on = on != 0;
  for (int k = 0; k < 4; ++k) {
    g_prof.ms[k] = 0.0;
    g_prof.n[k] = 0;
  }
  return 0;
}

int hops_set_option(const char *name, int value) {
  if (!name) return fail("hops_set_option: NULL name");
  if (std::string(name) == "stationary_stage") {
    g_stationary = value != 0;
    return 0;
  }
  if (std::string(name) == "chunk_traj") {
    if (value < 0) return fail("chunk_traj must be >= 0");
    g_chunk_traj = value;
    return 0;
  }
  if (std::string(name) == "nt_stream") {
    if (value < 0 || value > 4) return fail("nt_stream must be 0..4");
    g_nt_stream = value;
    return 0;
  }
  if (std::string(name) == "fuse_prologue") {
    g_fuse_linear = value != 0;
    return 0;
  }
  if (std::string(name) == "lds_stage") {
    g_lds_stage = value != 0;
    return 0;
  }
  if (std::string(name) == "vec_lanes") {
    if (value != 1 && value != 2 && value != 4) return fail("vec_lanes must be 1, 2 or 4");
    g_vec_lanes = value;
    return 0;
  }
  if (std::string(name) == "resident_rows") {
    if (value < 0) return fail("resident_rows must be >= 0");
    g_resident_rows = value;
    return 0;
  }
  if (std::string(name) == "graph_steps") {
    if (value < 0 || value > 4096) return fail("graph_steps must be 0..4096");
    g_graph_steps = value;
    return 0;
  }
  if (std::string(name) == "groups_q") {
    if (value < 0) return fail("groups_q must be >= 0");
    g_groups_q = value;
    return 0;
  }
  return fail("unknown option '%s'", name);
}

int hops_profile_read(double *ms_out, long long *n_out) {
  if (!ms_out || !n_out) return fail("hops_profile_read: NULL pointer");
  for (int k = 0; k < 4; ++k) {
    ms_out[k] = g_prof.ms[k];
    n_out[k] = g_prof.n[k];
  }
  return 0;
}

int hops_build_super(const hops_op_t *op, int precision, void *stream) {
  if (int rc = check_op(op)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const long nr = (long)op->n_state * op->n_hier;
  const unsigned g = (unsigned)((nr + 255) / 256);
  if (op->ell_w > 0) {
    if (precision == HOPS_PREC_FP64)
      k_build_ell<double><<<g, 256, 0, st>>>(*op);
    else if (precision == HOPS_PREC_MIXED)
      k_build_ell<float><<<g, 256, 0, st>>>(*op);
    else
      return fail("unknown precision %d", precision);
    HIP_TRY(hipGetLastError());
  }
  k_build_k0<<<(unsigned)((op->n_hier + 255) / 256), 256, 0, st>>>(*op);
  HIP_TRY(hipGetLastError());
  return 0;
}

int hops_derivative(const hops_op_t *op, const hops_batch_t *bt, const void *y_d, void *d_d, int tidx,
                    void *stream) {
  if (int rc = check_op(op)) return rc;
  if (!bt || !y_d || !d_d) return fail("hops_derivative: NULL argument");
  if (bt->precision != HOPS_PREC_FP64) return fail("hops_derivative needs an FP64 batch");
  if (tidx < 0 || tidx >= bt->n_t) return fail("noise index %d outside [0, %d)", tidx, bt->n_t);
  if (!bt->coefT_d || !bt->sigT_d || bt->coefT_stride < bt->B) return fail("stage coefficient table missing");
  hipStream_t st = (hipStream_t)stream;
  const int n = op->n_state;
  const size_t shm = sizeof(double) * (2 * (size_t)n + 8 * (size_t)op->n_act + 2 * (size_t)op->n_mode_abs + 16);
  if (shm > 160 * 1024) return fail("prologue LDS %zu B exceeds 160 KiB", shm);
  if (shm > 64 * 1024)
    HIP_TRY(hipFuncSetAttribute((const void *)k_prologue<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
  const cx<double> *y = reinterpret_cast<const cx<double> *>(y_d);
  if ((op->hd_d != nullptr) != (bt->hy_d != nullptr))
    return fail("dense system block needs its HY pool (and only it)");
  k_prologue<double><<<bt->B, 256, shm, st>>>(*op, *bt, -1, tidx, y, -1, nullptr);
  HIP_TRY(hipGetLastError());
  if (op->hd_d) {
    const int G = (bt->B + (1 << bt->lane_shift) - 1) >> bt->lane_shift;
    if (int rc = launch_dense_h<double>(*op, *bt, G, y, reinterpret_cast<cx<double> *>(bt->hy_d), st)) return rc;
  }
  k_stage<double><<<bt->n_tiles, 256, 0, st>>>(*op, *bt, -1, y, reinterpret_cast<cx<double> *>(d_d));
  HIP_TRY(hipGetLastError());
  return 0;
}

int hops_permute_c128(const void *src_d, void *dst_d, const int32_t *src_idx_d, const int32_t *dst_idx_d,
                      int n, double scale, void *stream) {
  if (n <= 0) return 0;
  if (!src_d || !dst_d || !src_idx_d || !dst_idx_d) return fail("hops_permute_c128: NULL argument");
  k_permute<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      reinterpret_cast<const double2 *>(src_d), reinterpret_cast<double2 *>(dst_d), src_idx_d, dst_idx_d, n,
      scale);
  HIP_TRY(hipGetLastError());
  return 0;
}

int hops_apply_state_op(const void *phi_in_d, void *phi_out_d, const void *op_d, int n_state, int n_hier, int B,
                        int lane_shift, void *stream) {
  if (!phi_in_d || !phi_out_d || !op_d) return fail("hops_apply_state_op: NULL argument");
  if (n_state <= 0 || n_hier <= 0 || B <= 0 || lane_shift < 0 || lane_shift > 6)
    return fail("hops_apply_state_op: bad sizes (n_state=%d n_hier=%d B=%d lane_shift=%d)", n_state, n_hier, B,
                lane_shift);
  if (phi_in_d == phi_out_d) return fail("hops_apply_state_op: input and output must differ");
  const long tot = (long)n_state * n_hier * B;
  k_state_op<<<(unsigned)((tot + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      reinterpret_cast<const double2 *>(phi_in_d), reinterpret_cast<double2 *>(phi_out_d),
      reinterpret_cast<const double2 *>(op_d), n_state, n_hier, B, lane_shift);
  HIP_TRY(hipGetLastError());
  return 0;
}

static int propagate_checked(const hops_op_t *op, const hops_batch_t *bt, int it0, int n_steps, int store0,
                             void *stream, void *ext_scr) {
  if (int rc = check_op(op)) return rc;
  if (!bt) return fail("batch is NULL");
  if (bt->B <= 0 || bt->tile_rows != 256 || bt->n_tiles <= 0) return fail("bad batch geometry");
  if (bt->lane_shift < 0 || bt->lane_shift > 4) return fail("lane_shift %d outside [0, 4]", bt->lane_shift);
  if (bt->eom < 0 || bt->eom > 2) return fail("unknown eom %d", bt->eom);
  const CoefLayout L = coef_layout(*op);
  if (bt->coef_stride < L.sig + 1) return fail("coef_stride %d < %d", bt->coef_stride, L.sig + 1);
  if (bt->ws_stride < 4 * op->n_mode_abs + 1) return fail("ws_stride too small");
  if (!bt->coefT_d || !bt->sigT_d || bt->coefT_stride < bt->B) return fail("stage coefficient table missing or too small");
  if (n_steps < 0) return fail("n_steps < 0");
  if (bt->precision != HOPS_PREC_FP64 && bt->precision != HOPS_PREC_MIXED)
    return fail("unknown precision %d", bt->precision);
  hipStream_t st = (hipStream_t)stream;
  auto run = [&](const hops_batch_t *b, int i0, int n, int s0) {
    return b->precision == HOPS_PREC_FP64 ? propagate_t<double>(op, b, i0, n, s0, st, ext_scr)
                                          : propagate_t<float>(op, b, i0, n, s0, st, ext_scr);
  };
  if (bt->sigma_raw && n_steps > 1) {
    // the raw first step alone, then the rest normalised as usual
    if (int rc = run(bt, it0, 1, store0)) return rc;
    hops_batch_t rest = *bt;
    rest.sigma_raw = 0;
    return run(&rest, it0 + 2 * bt->noise_step, n_steps - 1, store0 + 1);
  }
  return run(bt, it0, n_steps, store0);
}

int hops_propagate(const hops_op_t *op, const hops_batch_t *bt, int it0, int n_steps, int store0,
                   void *stream) {
  return propagate_checked(op, bt, it0, n_steps, store0, stream, nullptr);
}

int hops_propagate_noalloc(const hops_op_t *op, const hops_batch_t *bt, int it0, int n_steps, int store0,
                           void *stream) {
  void *scr = nullptr;
  if (int rc = noalloc_scratch(&scr)) return rc;
  return propagate_checked(op, bt, it0, n_steps, store0, stream, scr);
}


int hops_slots_step(const hops_op_t *ops_d, const hops_batch_t *bts_d, const int32_t *list_d, int n_list,
                    const uint32_t *tiles_d, int n_tiles, const int32_t *it_d, int precision,
                    int shm_doubles, void *stream) {
  if (int rc = slots_args_ok(ops_d, bts_d, list_d, n_list, tiles_d, n_tiles, (size_t)shm_doubles)) return rc;
  if (!it_d) return fail("hops_slots_step: NULL noise index array");
  const size_t shm = sizeof(double) * (size_t)shm_doubles;
  hipStream_t st = (hipStream_t)stream;
  if (precision == HOPS_PREC_FP64)
    return slots_step_t<double>(ops_d, bts_d, list_d, n_list, tiles_d, n_tiles, it_d, shm, false, st);
  if (precision == HOPS_PREC_MIXED)
    return slots_step_t<float>(ops_d, bts_d, list_d, n_list, tiles_d, n_tiles, it_d, shm, true, st);
  return fail("unknown precision %d", precision);
}

int hops_slots_pn(const hops_op_t *ops_d, const hops_batch_t *bts_d, const int32_t *list_d, int n_list,
                  const int64_t *off_d, void *out_d, double *nrm_d, int precision, int psi_only,
                  void *stream) {
  if (!ops_d || !bts_d || !list_d || !off_d || !out_d || (!psi_only && !nrm_d))
    return fail("hops_slots_pn: NULL argument");
  if (n_list <= 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  double2 *out = reinterpret_cast<double2 *>(out_d);
  if (precision == HOPS_PREC_FP64)
    k_pn_slots<double><<<n_list, 256, 0, st>>>(ops_d, bts_d, list_d, off_d, out, nrm_d, psi_only);
  else if (precision == HOPS_PREC_MIXED)
    k_pn_slots<float><<<n_list, 256, 0, st>>>(ops_d, bts_d, list_d, off_d, out, nrm_d, psi_only);
  else
    return fail("unknown precision %d", precision);
  HIP_TRY(hipGetLastError());
  return 0;
}

int hops_slots_errors(const hops_op_t *ops_d, const hops_batch_t *bts_d, const int32_t *list_d, int n_list,
                      const uint32_t *tiles_d, int n_tiles, const int32_t *it_d, int precision,
                      int shm_doubles, const int64_t *off_d, const void *pn_d, const int64_t *moff_d,
                      const uint8_t *mask_d, const int64_t *eoff_d, double *e_d, const double *dt_d,
                      double hbar, void *stream) {
  if (int rc = slots_args_ok(ops_d, bts_d, list_d, n_list, tiles_d, n_tiles, (size_t)shm_doubles)) return rc;
  if (!it_d || !off_d || !pn_d || !eoff_d || !e_d || !dt_d) return fail("hops_slots_errors: NULL argument");
  if ((moff_d == nullptr) != (mask_d == nullptr)) return fail("hops_slots_errors: mask without offsets");
  const size_t shm = sizeof(double) * (size_t)shm_doubles;
  hipStream_t st = (hipStream_t)stream;
  const double2 *pn = reinterpret_cast<const double2 *>(pn_d);
  const int mode = mask_d ? 1 : 0;
  if (precision == HOPS_PREC_FP64) {
    if (mask_d) {
      k_mask_slots<double><<<n_tiles, 256, 0, st>>>(ops_d, bts_d, tiles_d, off_d, pn, moff_d, mask_d);
      HIP_TRY(hipGetLastError());
    }
    if (int rc = slots_deriv_t<double>(ops_d, bts_d, list_d, n_list, tiles_d, n_tiles, it_d, shm, st)) return rc;
    k_err_slots<double><<<n_list, 256, 0, st>>>(ops_d, bts_d, list_d, off_d, pn, eoff_d, e_d, dt_d, hbar, mode,
                                                moff_d, mask_d);
  } else if (precision == HOPS_PREC_MIXED) {
    if (mask_d) {
      k_mask_slots<float><<<n_tiles, 256, 0, st>>>(ops_d, bts_d, tiles_d, off_d, pn, moff_d, mask_d);
      HIP_TRY(hipGetLastError());
    }
    if (int rc = slots_deriv_t<float>(ops_d, bts_d, list_d, n_list, tiles_d, n_tiles, it_d, shm, st)) return rc;
    k_err_slots<float><<<n_list, 256, 0, st>>>(ops_d, bts_d, list_d, off_d, pn, eoff_d, e_d, dt_d, hbar, mode,
                                               moff_d, mask_d);
  } else {
    return fail("unknown precision %d", precision);
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

int hops_slots_build(const hops_op_t *ops_d, const int32_t *list_d, int n_list, const uint32_t *tiles_d,
                     int n_tiles, int precision, void *stream) {
  if (!ops_d || !list_d || !tiles_d) return fail("hops_slots_build: NULL argument");
  if (n_list <= 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (n_tiles > 0) {
    if (precision == HOPS_PREC_FP64)
      k_build_ell_slots<double><<<n_tiles, 256, 0, st>>>(ops_d, tiles_d);
    else if (precision == HOPS_PREC_MIXED)
      k_build_ell_slots<float><<<n_tiles, 256, 0, st>>>(ops_d, tiles_d);
    else
      return fail("unknown precision %d", precision);
    HIP_TRY(hipGetLastError());
  }
  k_build_k0_slots<<<n_list, 256, 0, st>>>(ops_d, list_d);
  HIP_TRY(hipGetLastError());
  return 0;
}

int hops_slots_permute(const hops_slot_perm_t *perm_d, const uint32_t *tiles_d, int n_tiles, void *stream) {
  if (n_tiles <= 0) return 0;
  if (!perm_d || !tiles_d) return fail("hops_slots_permute: NULL argument");
  k_permute_slots<<<n_tiles, 256, 0, (hipStream_t)stream>>>(perm_d, tiles_d);
  HIP_TRY(hipGetLastError());
  return 0;
}

int hops_gather_u64(const uint64_t *const *src_d, uint64_t *out_d, int n, void *stream) {
  if (n <= 0) return 0;
  if (!src_d || !out_d) return fail("hops_gather_u64: NULL argument");
  k_gather_u64<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(src_d, out_d, n);
  HIP_TRY(hipGetLastError());
  return 0;
}

int hops_copy_segments(const hops_segment_t *segs_d, int n, void *stream) {
  if (n <= 0) return 0;
  if (!segs_d) return fail("hops_copy_segments: NULL argument");
  k_copy_segments<<<n, 256, 0, (hipStream_t)stream>>>(segs_d);
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // extern "C"
