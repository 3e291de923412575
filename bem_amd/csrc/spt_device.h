// Device-side maths for the MI355X tracing path: complete elliptic
// integrals of the complementary parameter, m = 0 ring kernels, block
// reductions. FP64 throughout; explicit fma() where fusion is wanted (the
// library is built with -ffp-contract=off so every other expression
// rounds like the NumPy reference).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "ellip_coeffs.h"
#include "log_coeffs.h"

namespace spt {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;
constexpr double kInv4Pi2 = 1.0 / (4.0 * kPi * kPi);
constexpr double kInv2Pi2 = 1.0 / (2.0 * kPi * kPi);
constexpr double kInvPi = 1.0 / kPi;

// Coefficient tables in constant memory: uniform across a wavefront, so
// they sit in scalar registers. Only K's pair is needed on device (E comes
// from K and dK/dp below), which keeps the table within the scalar file.
__constant__ double c_PK[11] = SPT_ELLIP_PK;
__constant__ double c_QK[11] = SPT_ELLIP_QK;

// Horner evaluation of a degree-10 polynomial and its derivative.
__device__ __forceinline__ void horner11_d(const double* c, double x, double& s, double& d) {
  d = c[10];
  s = fma(d, x, c[9]);
#pragma unroll
  for (int i = 8; i >= 0; --i) {
    d = fma(d, x, s);
    s = fma(s, x, c[i]);
  }
}

// Natural logarithm of a finite x >= 0 (x = 0 gives -inf), within 0.6 x
// 2^-52 relative: x = 2^e m, m in [sqrt(1/2), sqrt(2)), f = m - 1,
// s = f / (2 + f), ln m = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R(s^2))),
// R(z) = z T(z) with T fitted by gen_log_coeffs.py (the atanh reduction
// of Cody and Waite's logarithm). Only the small terms carry the rounding
// of s, so one Newton correction of the quotient suffices. About 30
// instructions against the ~75 of the library log, whose double-double
// steps this kernel does not need (K and E tolerate ~1 ulp in ln p).
__device__ __forceinline__ double log_pos(double x) {
  constexpr double T[7] = SPT_LOG_T;
  double m = __builtin_amdgcn_frexp_mant(x);
  int e = __builtin_amdgcn_frexp_exp(x);
  if (m < 0.70710678118654752440) {
    m = m + m;
    e = e - 1;
  }
  const double f = m - 1.0;
  const double d = 2.0 + f;
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  double s = f * r;
  s = fma(r, fma(-d, s, f), s);
  const double z = s * s;
  double t = T[6];
#pragma unroll
  for (int i = 5; i >= 0; --i) t = fma(t, z, T[i]);
  const double R = z * t;
  const double hf = 0.5 * (f * f);
  const double de = (double)e;
  const double v = fma(de, SPT_LOG_LN2_HI, -((hf - fma(s, hf + R, de * SPT_LOG_LN2_LO)) - f));
  return x == 0.0 ? -INFINITY : v;
}

// (1/c_i, ln c_i) of log_tab's 256 intervals; kernels that call log_tab
// copy them into LDS (fill_log_table) and pass the LDS copy.
__constant__ double c_logt_inv[SPT_LOGT_N] = SPT_LOGT_INVC;
__constant__ double c_logt_logc[SPT_LOGT_N] = SPT_LOGT_LOGC;

template <int T>
__device__ __forceinline__ void fill_log_table(double2* tab) {
  for (int k = threadIdx.x; k < SPT_LOGT_N; k += T) tab[k] = make_double2(c_logt_inv[k], c_logt_logc[k]);
}

// Natural logarithm of a normal x > 0 to an absolute 2^-52 max(1, |ln x|)
// (0.74 x 2^-52 measured), the accuracy the elliptic integrals below need
// of ln p (K and E are O(1) where ln p is small). x = 2^e m, m in [1/2, 1);
// the top 8 fraction bits of m pick the interval centre c_i, r = m/c_i - 1
// by one fma (|r| <= 2^-9), ln x = e ln2 + ln c_i + log1p(r) with log1p to
// the r^5 term (truncation below 2^-56). tab = (1/c_i, ln c_i), normally in
// LDS. 13 instructions against log_pos's ~30.
__device__ __forceinline__ double log_tab(double x, const double2* tab) {
  constexpr double Q[4] = SPT_LOGT_Q;
  const double m = __builtin_amdgcn_frexp_mant(x);
  const int e = __builtin_amdgcn_frexp_exp(x);
  const double2 t = tab[(__double2hiint(m) >> 12) & (SPT_LOGT_N - 1)];
  const double r = fma(m, t.x, -1.0);
  double q = fma(Q[3], r, Q[2]);
  q = fma(q, r, Q[1]);
  q = fma(q, r, Q[0]);
  return fma((double)e, SPT_LOG_LN2, t.y + fma(r * r, q, r));
}

// K(1-p), E(1-p) for p in [0, 1], given lp = ln p.
//   K(1-p) = PK(p) - ln(p) QK(p)  (logarithmic form of DLMF 19.12.1,
//   coefficients fitted by gen_ellip_coeffs.py, ~1.2 ulp);
//   E from Legendre's derivative relation dK/dm = (E - (1-m)K)/(2m(1-m))
//   (DLMF 19.4.1) with m = 1 - p:  E = p K - 2 (1-p) p dK/dp,
//   p dK/dp = p PK'(p) - QK(p) - p ln(p) QK'(p)  (~1.6 ulp).
__device__ __forceinline__ void ellip_ke_lp(double p, double lp, double& K, double& E) {
  double P, dP, Q, dQ;
  horner11_d(c_PK, p, P, dP);
  horner11_d(c_QK, p, Q, dQ);
  K = fma(-lp, Q, P);
  const double pkp = fma(p, dP, -Q) - (p * lp) * dQ;
  E = fma(p, K, -2.0 * (1.0 - p) * pkp);
}

__device__ __forceinline__ void ellip_ke_c(double p, double& K, double& E) { ellip_ke_lp(p, log_pos(p), K, E); }

__device__ __forceinline__ double ellip_k_c(double p) {
  const double lp = log_pos(p);
  double s = c_PK[10], q = c_QK[10];
#pragma unroll
  for (int i = 9; i >= 0; --i) {
    s = fma(s, p, c_PK[i]);
    q = fma(q, p, c_QK[i]);
  }
  return fma(-lp, q, s);
}

// 1/x to ~1 ulp: hardware reciprocal estimate r (relative error ~2^-23),
// e = 1 - x r, then one third-order step 1/x = r (1 + e + e^2) + O(e^3).
__device__ __forceinline__ double rcp64(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  const double e = fma(-x, r, 1.0);
  return fma(r, fma(e, e, e), r);
}

// 1/sqrt(x) to ~1 ulp: hardware estimate y (relative error ~2^-23),
// e = 1 - x y^2, then one third-order step
// 1/sqrt(x) = y (1 + e/2 + 3 e^2/8) + O(e^3).
__device__ __forceinline__ double rsqrt64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-(x * y), y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Near-axis Taylor terms of the ring field (reference
// bem/axisymmetric.py:262-264), full constants included.
// The half-integer powers are ad2^k sqrt(ad2) (a few ulps from pow).
__device__ __forceinline__ double2 ring_field_axis(double rho, double z, double a) {
  const double ad2 = a * a + z * z;
  const double p15 = ad2 * sqrt(ad2), p25 = ad2 * p15, p35 = ad2 * p25;
  double2 f;
  f.x = rho * (2 * z * z - a * a) / (8 * kPi * p25);
  f.y = z / (4 * kPi * p15) + 3 * rho * rho * z * (3 * a * a - 2 * z * z) / (16 * kPi * p35);
  return f;
}

// Weighted ring-field accumulation for one source ring (radius a, axial
// offset zz = z_point - z_ring, weight w = charge). The main-branch terms
// are accumulated without their per-point constants:
//   er_main += w (K - (a^2 - rho^2 + zz^2) E / d2) / sqrt(D)   [x 1/(4 pi^2 rho) later]
//   ez_main += w zz E / (sqrt(D) d2)                           [x 1/(2 pi^2) later]
// Near-axis nodes (rho^2 < 1e-10 (a^2 + zz^2)) go to er_ax/ez_ax with
// their constants (reference bem/axisymmetric.py:248-265).
struct RingAcc {
  double er_main, ez_main, er_ax, ez_ax;
  __device__ void zero() { er_main = ez_main = er_ax = ez_ax = 0.0; }
  __device__ void add(const RingAcc& o) {
    er_main += o.er_main;
    ez_main += o.ez_main;
    er_ax += o.er_ax;
    ez_ax += o.ez_ax;
  }
};

// Main-branch term of ring_accum for a node known not to be near-axis
// (a2 = a^2). lt: log_tab's table (p > 0 here: the gather never places a
// node on the point).
__device__ __forceinline__ void ring_accum_main(double rho, double rho2, double zz, double a, double a2, double w,
                                                RingAcc& acc, const double2* lt) {
  const double zz2 = zz * zz;
  const double sp = rho + a, sm = rho - a;
  const double D = fma(sp, sp, zz2);
  const double d2 = fma(sm, sm, zz2);
  const double rD = rsqrt64(D);
  const double p = d2 * (rD * rD);
  double K, E;
  const double pc = fmin(p, 1.0);
  ellip_ke_lp(pc, log_tab(pc, lt), K, E);
  const double u = E * rcp64(d2);  // E / d2
  const double c = (a2 - rho2) + zz2;
  const double wr = w * rD;
  acc.er_main = fma(wr, fma(-c, u, K), acc.er_main);
  acc.ez_main = fma(wr * zz, u, acc.ez_main);
}

__device__ __forceinline__ void ring_accum(double rho, double rho2, double zz, double a, double w, RingAcc& acc,
                                           const double2* lt) {
  const double a2 = a * a;
  if (rho2 < 1e-10 * (a2 + zz * zz)) {
    const double2 f = ring_field_axis(rho, zz, a);
    acc.er_ax = fma(w, f.x, acc.er_ax);
    acc.ez_ax = fma(w, f.y, acc.ez_ax);
    return;
  }
  ring_accum_main(rho, rho2, zz, a, a2, w, acc, lt);
}

// Can any node of the segment (p0r, p0z) + t (abr, abz), t in [0, 1],
// take ring_accum's near-axis branch for a field point at (rho, z)? Its
// nodes have a <= max(p0r, p1r) and |z - z_node| <= max(|z - p0z|,
// |z - p1z|), so no node does when rho^2 >= 1e-10 (a_max^2 + zz_max^2);
// the 1 % margin covers the rounding of the node table and of the test.
__device__ __forceinline__ bool ring_axis_possible(double rho2, double z, double p0r, double p0z, double abr,
                                                   double abz) {
  const double amax = fmax(fabs(p0r), fabs(p0r + abr));
  const double zmax = fmax(fabs(z - p0z), fabs(z - (p0z + abz)));
  return rho2 < 1.01e-10 * (amax * amax + zmax * zmax);
}

__device__ __forceinline__ void ring_finish(const RingAcc& acc, double rho, double& er, double& ez) {
  er = (rho > 0.0 ? acc.er_main * kInv4Pi2 / rho : 0.0) + acc.er_ax;
  ez = acc.ez_main * kInv2Pi2 + acc.ez_ax;
}

// Potential of a unit-total-charge ring, G = 1/(4 pi r) (reference
// bem/axisymmetric.py:221-229).
__device__ __forceinline__ double ring_potential(double rho, double z, double a) {
  const double sp = rho + a, sm = rho - a;
  const double D = sp * sp + z * z;
  const double p = (sm * sm + z * z) / D;
  return ellip_k_c(fmin(p, 1.0)) / (2.0 * kPi * kPi * sqrt(D));
}

// ---------------------------------------------------- FP32 (mixed) path
// Degree-6 FP32 pair for K (same logarithmic form), E from the same
// Legendre derivative relation; ~1 and ~1.5 float ulp.
__constant__ float c_PK32[7] = SPT_ELLIP_PK32;
__constant__ float c_QK32[7] = SPT_ELLIP_QK32;

__device__ __forceinline__ void horner7f_d(const float* c, float x, float& s, float& d) {
  s = c[6];
  d = 0.0f;
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    d = fmaf(d, x, s);
    s = fmaf(s, x, c[i]);
  }
}

__device__ __forceinline__ void ellip_ke_c32(float p, float& K, float& E) {
  const float lp = __logf(p);
  float P, dP, Q, dQ;
  horner7f_d(c_PK32, p, P, dP);
  horner7f_d(c_QK32, p, Q, dQ);
  K = fmaf(-lp, Q, P);
  const float pkp = fmaf(p, dP, -Q) - (p * lp) * dQ;
  E = fmaf(p, K, -2.0f * (1.0f - p) * pkp);
}

// Kahan-compensated FP32 accumulator: the true sum is s - c.
struct Kahan32 {
  float s, c;
  __device__ void zero() { s = c = 0.0f; }
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

struct RingAcc32 {
  Kahan32 er, ez;
  __device__ void zero() {
    er.zero();
    ez.zero();
  }
};

// FP32 version of ring_accum for one far node. rho, a and the axial offset
// arrive in FP64 and their sums/differences are formed in FP64 before
// rounding, so the 1e-5 m absolute coordinates do not cost relative
// accuracy in the 1e-7 m separations. Near-axis nodes use the FP64 Taylor
// branch (into the FP64 accumulator).
__device__ __forceinline__ void ring_accum32(double rho, double rho2, double zz_d, double a, float w, RingAcc32& acc,
                                             RingAcc& acc64) {
  const float zz = (float)zz_d;
  const float zz2 = zz * zz;
  const float af = (float)a;
  if ((float)rho2 < 1e-10f * (af * af + zz2)) {
    const double2 f = ring_field_axis(rho, zz_d, a);
    acc64.er_ax = fma((double)w, f.x, acc64.er_ax);
    acc64.ez_ax = fma((double)w, f.y, acc64.ez_ax);
    return;
  }
  const float sp = (float)(rho + a), sm = (float)(rho - a);
  const float D = fmaf(sp, sp, zz2);
  const float d2 = fmaf(sm, sm, zz2);
  const float rD = __builtin_amdgcn_rsqf(D);
  const float p = fminf(d2 * (rD * rD), 1.0f);
  float K, E;
  ellip_ke_c32(p, K, E);
  const float inv_d2 = __builtin_amdgcn_rcpf(d2);
  const float c = fmaf(-sm, sp, zz2);  // a^2 - rho^2 + zz^2
  const float wr = w * rD;
  acc.er.add(wr * fmaf(-c * E, inv_d2, K));
  acc.ez.add((wr * zz) * (E * inv_d2));
}

// ---- block reductions (deterministic: fixed tree over thread index) ----
// One-wavefront blocks (T = 64) reduce with shuffles in the same pairing
// (lane i takes lane i + s, s = 32 .. 1, the result is lane 0's), so the
// sum is bit-identical to the LDS tree, without its seven LDS round trips;
// the barriers at both ends keep the tree's ordering of other LDS traffic.
template <int T>
__device__ double block_sum(double v, double* red) {
  if constexpr (T == 64) {
    __syncthreads();
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v = v + __shfl_down(v, s, 64);
    const double r = __shfl(v, 0, 64);
    __syncthreads();
    return r;
  }
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
#pragma unroll
  for (int s = T / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] = red[tid] + red[tid + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

template <int T>
__device__ double block_min(double v, double* red) {
  if constexpr (T == 64) {
    __syncthreads();
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v = fmin(v, __shfl_down(v, s, 64));
    const double r = __shfl(v, 0, 64);
    __syncthreads();
    return r;
  }
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
#pragma unroll
  for (int s = T / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fmin(red[tid], red[tid + s]);
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

}  // namespace spt
