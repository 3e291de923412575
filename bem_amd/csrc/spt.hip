// libspt_hip: MI355X (gfx950) kernels and C ABI for the
// specific_particle_tracer tracing path. Declarations: include/spt.h.
//
// Layout in HBM
//   BEM node table   double4[n_seg * n_sub * 4] = (rho_q, z_q, area_q*sigma_q, 0),
//                    segment-major so a lane walks one segment contiguously.
//   segment table    double[n_seg * 12] (p0, tangent, |t|^2, length, unit, normal).
//   particle state   caller's (n, 3) pos/vel, CSR group offsets.
//   stage buffer     library scratch double[7 stages][6 comps][n] (coalesced
//                    across the threads of a group).
//
// Execution
//   spt_trace launches one persistent kernel. Each workgroup pulls a group
//   index from a device queue and integrates that group to t_max with its
//   own t and h: stage positions in LDS, external field by a block-
//   cooperative BEM gather (or closed form), in-group pair forces from LDS,
//   RMS error norm by a fixed-order block reduction, accept/reject, birth
//   checkpoints, kill and screen crossings, all on device.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <algorithm>
#include <vector>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <string>

#include "../../include/spt.h"
#include "spt_device.h"

using namespace spt;

// Gauss-Legendre nodes/weights exactly as numpy.polynomial.legendre.leggauss
// returns them (the reference's rules, bem/axisymmetric.py:94,98).
constexpr double c_gx4[4] = {-0.8611363115940526, -0.33998104358485626, 0.33998104358485626, 0.8611363115940526};
constexpr double c_gw4[4] = {0.3478548451374537, 0.6521451548625462, 0.6521451548625462, 0.3478548451374537};
constexpr double c_gx8[8] = {-0.9602898564975362, -0.7966664774136267, -0.525532409916329, -0.18343464249564978,
                                0.18343464249564978, 0.525532409916329, 0.7966664774136267, 0.9602898564975362};
constexpr double c_gw8[8] = {0.10122853629037669, 0.22238103445337434, 0.31370664587788705, 0.36268378337836177,
                                0.36268378337836177, 0.31370664587788705, 0.22238103445337434, 0.10122853629037669};

// Dormand-Prince 5(4) tableau (Dormand & Prince 1980), the pair the
// reference integrates with (batched_rk45.py:30-50).
__constant__ double c_A[6][5] = {
    {0, 0, 0, 0, 0},
    {1.0 / 5, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656},
};
__constant__ double c_B[6] = {35.0 / 384, 0.0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84};
__constant__ double c_E[7] = {71.0 / 57600, 0.0, -71.0 / 16695, 71.0 / 1920, -17253.0 / 339200, 22.0 / 525, -1.0 / 40};

// 1/(4 pi eps0) with CODATA 2022 eps0, folded in the same operation order
// as the reference (constants.py:14): 8987551786.170797.
constexpr double kVacuumPermittivity = 8.8541878188e-12;
constexpr double kCoulomb = 1.0 / (4.0 * kPi * kVacuumPermittivity);
constexpr double kLight = 299792458.0;

// A group stalls when an attempt is rejected and the next step is not above
// kStallFraction |t| (zero and NaN steps included): the controller can no
// longer advance it. Rejections shrink h by at least min_factor, so a
// stuck group reaches this in a bounded number of attempts instead of
// spinning to max_steps (the reference's only exit,
// batched_rk45.py:320-321).
constexpr double kStallFraction = 1e-14;
__device__ __forceinline__ bool step_stalled(bool accepted, double h_next, double t) {
  return !accepted && !(h_next > kStallFraction * fabs(t));
}

// ---------------------------------------------------------------- tables
enum { S_P0R = 0, S_P0Z, S_ABR, S_ABZ, S_ABAB, S_L, S_UR, S_UZ, S_NR, S_NZ, SEG_STRIDE = 12 };

struct BemView {
  int n_seg;
  int nq;                 // nodes per segment (n_sub * 4)
  const double4* nodes;   // n_seg * nq
  const double* segs;     // n_seg * SEG_STRIDE
  const double* sigma;    // n_seg + 1
  // Distance-adaptive far rule (nullptr: the reference's fixed rule).
  // Per segment kCoarsePer nodes: Gauss-4 over the whole segment, Gauss-2,
  // Gauss-8; a segment at distance D >= L from the point takes Gauss-8
  // below k4 L, Gauss-4 below k2 L and Gauss-2 beyond, instead of its nq
  // sub-panel nodes.
  const double4* coarse;
  double k4, k2;
  // Exterior multipole expansion of the surface charge about the origin
  // (nullptr: off): per term l, (m_l, (2l+1)/(l+1), l/(l+1), 0); used for
  // points at r^2 >= mp_rmin2 instead of the segment sums.
  const double4* mp;
  int mp_n;
  double mp_rref, mp_rmin2;
  // Axis-centred expansions of runs of consecutive segments (far-rule
  // gathers only; nullptr: off): per cluster (z_c, r_c, (kappa r_c)^2, 0),
  // its segments cl_seg[c] .. cl_seg[c + 1] - 1, moments cl_mom[c][l]; a
  // point at (kappa r_c)^2 or more from (0, z_c) takes the cluster's field
  // from the expansion, truncated where the tail bound drops below cl_tol.
  const double4* cl;
  const int* cl_seg;
  const int* seg_cl;        // per segment: its cluster
  const double* cl_mom;     // cl_n x cl_terms
  const double2* cl_coef;   // per term: ((2l+1)/(l+1), l/(l+1))
  int cl_n, cl_terms;
  double cl_tol;
  double near_w;            // near-rule interval width of far-rule gathers (the reference's: 0.5)
  // Chebyshev tiles (far-rule gathers only; tl_tree nullptr: off): a
  // quadtree over the square [tl_x0, tl_x0 + tl_size) x [tl_z0, tl_z0 +
  // tl_size) of (rho, z). tl_tree[k] >= 0: index of node k's first child
  // (children in the order low/high rho, then low/high z); -1: no tile;
  // <= -2: tile -2 - tl_tree[k]. A point in tile t sums its window's
  // segments tl_win[t].x .. tl_win[t].y - 1 with the far-rule gather's
  // rules and takes every other segment's fixed-rule field from the tile's
  // table: kTileP x kTileP Chebyshev coefficients of E_rho, then of E_z,
  // over the cell (tl_coef + t * kTileCoef). tl_cell[t]: tile t's cell
  // (low rho, low z, side) as the walk leaves them; tiles tl_nwin .. tl_n - 1
  // have empty windows.
  const int* tl_tree;
  const int2* tl_win;
  const double* tl_coef;
  const double4* tl_cell;
  int tl_n, tl_nwin;
  double tl_x0, tl_z0, tl_size;
};
constexpr int kPackThreads = 256;  // threads of a block of packed 16-lane groups (lockstep accel and end kernels)
constexpr int kTileP = 12;                         // Chebyshev degree + 1 per direction
constexpr int kTileCoef = 2 * kTileP * kTileP;     // doubles per tile
constexpr int kTileMaxDepth = 24;
constexpr int kTileMaxTiles = 6144;
constexpr int kCoarsePer = 14;
constexpr int kMultipoleMaxTerms = 1024;
constexpr double kMultipoleTail = 1e-12;  // multipole_pert's truncation bound (relative to max |m_l|)
constexpr int kMaxClusters = 64;
constexpr int kClusterMaxTerms = 256;
constexpr double kRefNearWidth = 0.5;

// Segment geometry (reference bem/axisymmetric.py:329-338,490-498).
__global__ void seg_table_kernel(const double* prof, int n_seg, double* segs) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_seg) return;
  const double p0r = prof[2 * j], p0z = prof[2 * j + 1];
  const double tr = prof[2 * j + 2] - p0r, tz = prof[2 * j + 3] - p0z;
  const double L = sqrt(tr * tr + tz * tz);
  const double nn = sqrt((-tz) * (-tz) + tr * tr);
  double* g = segs + (size_t)j * SEG_STRIDE;
  g[S_P0R] = p0r;
  g[S_P0Z] = p0z;
  g[S_ABR] = tr;
  g[S_ABZ] = tz;
  g[S_ABAB] = tr * tr + tz * tz;
  g[S_L] = L;
  g[S_UR] = tr / L;
  g[S_UZ] = tz / L;
  g[S_NR] = -tz / nn;
  g[S_NZ] = tr / nn;
  g[10] = 0.0;
  g[11] = 0.0;
}

// Fixed quadrature node table (reference bem/axisymmetric.py:305-326):
// (rho_q, z_q, weight, rho_q^2), weight = area_q * sigma_q with
// area_q = 2 pi rho_q (L / n_sub) w_g / 2.
__global__ void node_table_kernel(const double* prof, const double* sigma, int n_seg, int n_sub, double4* nodes) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = n_sub * 4;
  if (k >= n_seg * per) return;
  const int j = k / per, r = k - j * per, sub = r / 4, gq = r - sub * 4;
  const double p0r = prof[2 * j], p0z = prof[2 * j + 1];
  const double p1r = prof[2 * j + 2], p1z = prof[2 * j + 3];
  const double dr = p1r - p0r, dz = p1z - p0z;
  const double L = sqrt(dr * dr + dz * dz);
  const double ta = (double)sub / (double)n_sub, tb = (double)(sub + 1) / (double)n_sub;
  const double t = ta + 0.5 * (tb - ta) * (c_gx4[gq] + 1.0);
  const double rho = p0r + t * (p1r - p0r);
  const double z = p0z + t * (p1z - p0z);
  const double seg_len = (L / n_sub) * 0.5 * c_gw4[gq];
  const double area = 2.0 * kPi * rho * seg_len;
  const double sq = sigma[j] + t * (sigma[j + 1] - sigma[j]);
  nodes[k] = make_double4(rho, z, area * sq, rho * rho);
}

// Coarse far-rule nodes per segment (no counterpart in the reference, whose
// far rule is the fixed sub-panel table above): Gauss-4 over the whole
// segment (nodes 0-3), Gauss-2 (4-5), Gauss-8 (6-13), as (rho, z,
// area * sigma, rho^2) with sigma linear along the segment. Gauss-n over a
// segment of length L at distance D converges like (4D/L)^(-2n): with
// Gauss-8 at L <= D < 2L, Gauss-4 at 2L <= D < 16L and Gauss-2 beyond, the
// field changes by < 3.5e-8 of |E| on the 400-panel post against the fixed
// rule.
__global__ void coarse_table_kernel(const double* prof, const double* sigma, int n_seg, double4* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_seg * kCoarsePer) return;
  const int j = k / kCoarsePer, r = k - j * kCoarsePer;
  const double x = r < 4 ? c_gx4[r] : (r < 6 ? (r == 4 ? -0.5773502691896257645 : 0.5773502691896257645) : c_gx8[r - 6]);
  const double w = r < 4 ? c_gw4[r] : (r < 6 ? 1.0 : c_gw8[r - 6]);
  const double p0r = prof[2 * j], p0z = prof[2 * j + 1];
  const double p1r = prof[2 * j + 2], p1z = prof[2 * j + 3];
  const double dr = p1r - p0r, dz = p1z - p0z;
  const double L = sqrt(dr * dr + dz * dz);
  const double t = 0.5 * (x + 1.0);
  const double rho = p0r + t * dr;
  const double area = 2.0 * kPi * rho * (L * 0.5 * w);
  const double sq = sigma[j] + t * (sigma[j + 1] - sigma[j]);
  out[k] = make_double4(rho, p0z + t * dz, area * sq, rho * rho);
}

// ------------------------------------------------------- BEM field gather
// Minimum waves per SIMD requested for the field-gather kernels (caps their
// VGPR budget at 512 / this).
#ifndef SPT_FIELD_MIN_WAVES
#define SPT_FIELD_MIN_WAVES 3
#endif
// Near segment (closer than its own length): s = s0 + d sinh(v), Gauss-8 in
// intervals of width <= 0.5 over [asinh(-s0/d), asinh((L-s0)/d)], density
// linear between the segment's nodal values (reference
// bem/axisymmetric.py:571-600). sinh and cosh of a node come from one
// exp (their absolute error, ~1e-16 cosh v, moves the node by ~1e-16 d).
struct NearRule {
  double d, s0, lo, step, half;
  int count;
};

__device__ __forceinline__ NearRule near_rule(const double* g, double proj, double dist,
                                              double width = kRefNearWidth) {
  NearRule r;
  const double L = g[S_L];
  r.d = fmax(dist, 1e-15 * L);
  r.s0 = proj;
  r.lo = asinh(-r.s0 / r.d);
  const double hi = asinh((L - r.s0) / r.d);
  r.count = max(1, (int)ceil((hi - r.lo) / width));
  r.step = (hi - r.lo) / r.count;
  r.half = 0.5 * (hi - r.lo) / r.count;
  return r;
}

// near_rule for far-rule gathers: the same substitution and interval
// rule, with asinh x = sign(x) ln(|x| + sqrt(x^2 + 1)) from the LDS log
// table (absolute error ~2e-16 in v, so the integration limits land within
// ~2e-16 |s - s0| of the segment's ends) and reciprocals (rcp64, ~1 ulp)
// instead of divisions.
__device__ __forceinline__ NearRule near_rule_fast(const double* g, double proj, double dist, double width,
                                                   const double2* lt) {
  NearRule r;
  const double L = g[S_L];
  r.d = fmax(dist, 1e-15 * L);
  r.s0 = proj;
  const double id = rcp64(r.d);
  auto ash = [&](double x) {
    const double a = fabs(x);
    return copysign(log_tab(a + sqrt(fma(a, a, 1.0)), lt), x);
  };
  r.lo = ash(-r.s0 * id);
  const double hi = ash((L - r.s0) * id);
  r.count = max(1, (int)ceil((hi - r.lo) * rcp64(width)));
  r.step = (hi - r.lo) * rcp64((double)r.count);
  r.half = 0.5 * r.step;
  return r;
}

// Per-segment constants of the near rule's nodes.
struct NearSeg {
  double p0r, p0z, ur, uz, inv_L, sig0, dsig;
};

__device__ __forceinline__ NearSeg near_seg(const double* g, double sig0, double sig1) {
  return {g[S_P0R], g[S_P0Z], g[S_UR], g[S_UZ], 1.0 / g[S_L], sig0, sig1 - sig0};
}

// Gauss node i (interval i / 8, node i % 8) of a segment's near rule. AXIS:
// the segment may meet the near-axis branch (ring_axis_possible), so each
// node takes ring_accum's per-node test; otherwise the main term directly
// (no Taylor-branch code in the loop).
template <bool AXIS>
__device__ __forceinline__ void near_node(double re, double re2, double ze, const NearSeg& sg, const NearRule& r,
                                          int i, RingAcc& acc, const double2* lt) {
  const int k = i >> 3, q = i & 7;
  const double v = r.lo + (k + 0.5 * (c_gx8[q] + 1.0)) * r.step;
  const double ev = exp(v), iv = rcp64(ev);
  const double sh = 0.5 * (ev - iv), ch = 0.5 * (ev + iv);
  const double s = r.s0 + r.d * sh;
  const double sr = sg.p0r + s * sg.ur;
  const double sz = sg.p0z + s * sg.uz;
  const double dens = sg.sig0 + (s * sg.inv_L) * sg.dsig;
  const double w = dens * (2 * kPi * sr) * r.d * ch * r.half * c_gw8[q];
  if (AXIS)
    ring_accum(re, re2, ze - sz, sr, w, acc, lt);
  else
    ring_accum_main(re, re2, ze - sz, sr, sr * sr, w, acc, lt);
}

// Nodes i0, i0 + stride, ... of segment g's near rule.
__device__ __forceinline__ void near_nodes(double re, double re2, double ze, const double* g, double sig0, double sig1,
                                           const NearRule& r, int i0, int stride, RingAcc& acc, const double2* lt) {
  const int nodes = 8 * r.count;
  if (i0 >= nodes) return;  // wide point blocks: most slices hold no node of a given segment
  const NearSeg sg = near_seg(g, sig0, sig1);
  if (ring_axis_possible(re2, ze, g[S_P0R], g[S_P0Z], g[S_ABR], g[S_ABZ])) {
    for (int i = i0; i < nodes; i += stride) near_node<true>(re, re2, ze, sg, r, i, acc, lt);
  } else {
    for (int i = i0; i < nodes; i += stride) near_node<false>(re, re2, ze, sg, r, i, acc, lt);
  }
}

// Pair form of a segment's near-rule nodes, for far-rule gathers: the
// Gauss-8 points are symmetric, so nodes p and 7 - p of interval k sit at
// v_k -+ |g_p| half about the interval's midpoint v_k, and
// exp(v) = exp(v_k) exp(-+|g_p| half): one exp per interval and one per
// pair instead of one per node. The products round within a few ulp of
// exp(v), which moves a node by ~1e-16 of its distance. Items m = 4k + p
// (interval k, pair p of nodes p and 7 - p); this thread takes m = i0,
// i0 + stride, ... with stride 2 or 4, so the same one or two pairs in
// every interval.
template <bool AXIS>
__device__ __forceinline__ void near_pair_nodes(double re, double re2, double ze, const NearSeg& sg, const NearRule& r,
                                                int i0, int stride, RingAcc& acc, const double2* lt) {
  const int np = 4 / stride;
  double e[2], ei[2], wk[2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    if (a < np) {
      const int p = i0 + a * stride;
      e[a] = exp(c_gx8[p] * r.half);  // g_p < 0: node p is below the midpoint
      ei[a] = rcp64(e[a]);
      wk[a] = (2 * kPi) * r.d * r.half * c_gw8[p];
    }
  }
  auto node = [&](double ev, double iv, double wq) {
    const double sh = 0.5 * (ev - iv), ch = 0.5 * (ev + iv);
    const double s = r.s0 + r.d * sh;
    const double sr = sg.p0r + s * sg.ur;
    const double sz = sg.p0z + s * sg.uz;
    const double dens = sg.sig0 + (s * sg.inv_L) * sg.dsig;
    const double w = dens * sr * ch * wq;
    if (AXIS)
      ring_accum(re, re2, ze - sz, sr, w, acc, lt);
    else
      ring_accum_main(re, re2, ze - sz, sr, sr * sr, w, acc, lt);
  };
  for (int k = 0; k < r.count; ++k) {
    const double c = exp(r.lo + (k + 0.5) * r.step), ci = rcp64(c);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      if (a < np) {
        node(c * e[a], ci * ei[a], wk[a]);
        node(c * ei[a], ci * e[a], wk[a]);
      }
    }
  }
}

__device__ __forceinline__ void near_pairs(double re, double re2, double ze, const double* g, double sig0, double sig1,
                                           const NearRule& r, int i0, int stride, RingAcc& acc, const double2* lt) {
  const NearSeg sg = near_seg(g, sig0, sig1);
  if (ring_axis_possible(re2, ze, g[S_P0R], g[S_P0Z], g[S_ABR], g[S_ABZ]))
    near_pair_nodes<true>(re, re2, ze, sg, r, i0, stride, acc, lt);
  else
    near_pair_nodes<false>(re, re2, ze, sg, r, i0, stride, acc, lt);
}

// Whole near segment on one thread; its sum comes back by value (the
// caller adds it), so the caller's accumulator never goes to memory.
__device__ __noinline__ RingAcc near_segment(double re, double re2, double ze, const double* g, double sig0,
                                             double sig1, double proj, double dist, const double2* lt,
                                             double width) {
  const NearRule r = near_rule(g, proj, dist, width);
  RingAcc acc;
  acc.zero();
  near_nodes(re, re2, ze, g, sig0, sig1, r, 0, 1, acc, lt);
  return acc;
}

// Projection of the point (re, ze) onto segment g, clamped to it, and the
// distance to that point.
template <class G>
__device__ __forceinline__ void segment_proj(double re, double ze, G g, double& proj, double& dist) {
  const double L = g[S_L], ur = g[S_UR], uz = g[S_UZ];
  const double dr = re - g[S_P0R], dz = ze - g[S_P0Z];
  const double along = dr * ur + dz * uz;
  proj = fmin(fmax(along, 0.0), L);
  const double qr = dr - proj * ur, qz = dz - proj * uz;
  dist = sqrt(qr * qr + qz * qz);
}

constexpr int kNearSlots = 7;  // near segments a slice hands to the point's cooperative pass

template <int T, int NMAX>
struct BemShared {
  double2 logt[SPT_LOGT_N];              // log_tab's table
  unsigned long long nmask[T / 64];      // per wavefront: lanes that handed over near segments
  double re[NMAX], ze[NMAX], rho[NMAX], cph[NMAX], sph[NMAX];
  union {
    struct {
      double bd2[T];
      int bseg[T];
    } a;
    struct {
      int cnt[T];
      int seg[kNearSlots][T];
    } n;
    double acc[4][T];
  } u;
};

// Node and segment tables read through the constant address space: where
// the index is uniform over the wavefront these become scalar loads.
typedef const __attribute__((address_space(4))) double4* cnode_ptr;
typedef const __attribute__((address_space(4))) double* cseg_ptr;

// Node q of a table read through cnode_ptr, component by component.
__device__ __forceinline__ double4 load_node(cnode_ptr nd, int q) {
  const cseg_ptr c = (cseg_ptr)(nd + q);
  return make_double4(c[0], c[1], c[2], c[3]);
}
__device__ __forceinline__ double4 load_node(const double4* nd, int q) { return nd[q]; }

// BEM perturbation field from the exterior multipole expansion of the
// surface charge (no counterpart in the reference, which sums the segments
// at every point): phi = (1/4 pi) sum_l m_l (r_ref/r)^l P_l(cos theta) / r
// about the origin, valid outside the sphere r = r_ref holding every
// quadrature node, so E_r = (1/4 pi r^2) sum m_l (l+1) q^l P_l and
// E_theta = (sin theta / 4 pi r^2) sum m_l q^l P_l' (q = r_ref / r), with
// P_{l+1} = a_l c P_l - b_l P_{l-1} and P'_{l+1} = P'_{l-1} + (2l+1) P_l.
// Terms stop after the first l whose tail bound sum_{k > l} (k+1) q^k is
// <= kMultipoleTail (|P_k| <= 1 and |sin t P'_k| <= k + 1, so the dropped
// terms of either component are below kMultipoleTail max_k |m_k| / 4 pi r^2).
// Returns false (nothing written) where r^2 < mp_rmin2.
__device__ __forceinline__ bool multipole_zone(const BemView& B, double x, double y, double z) {
  const double rho2 = x * x + y * y;
  return rho2 + z * z >= B.mp_rmin2;
}
__device__ __forceinline__ bool multipole_pert(const BemView& B, double x, double y, double z, double& ox, double& oy,
                                               double& oz) {
  const double rho2 = x * x + y * y;
  const double r2 = rho2 + z * z;
  if (!(r2 >= B.mp_rmin2)) return false;  // multipole_zone, the same arithmetic
  const double r = sqrt(r2), ir = 1.0 / r, rho = sqrt(rho2);
  const double c = z * ir, sn = rho * ir, q = B.mp_rref * ir;
  const double omq = 1.0 - q;
  const double tolw = kMultipoleTail * omq * omq;
  double pl = 1.0, plm = 0.0, dl = 0.0, dlm = 0.0, qp = 1.0, sa = 0.0, sb = 0.0;
  double lp1 = 1.0, tl1 = 1.0;  // l + 1 and 2l + 1, exact
  const cseg_ptr t = (cseg_ptr)B.mp;
  for (int l = 0; l < B.mp_n; ++l) {
    const double m = t[4 * l], a = t[4 * l + 1], b = t[4 * l + 2];
    const double mq = m * qp;
    sa = fma(mq * lp1, pl, sa);
    sb = fma(mq, dl, sb);
    qp *= q;
    if (qp * fma(-lp1, q, lp1 + 1.0) <= tolw) break;
    const double pn = fma(a * c, pl, -b * plm);
    const double dn = fma(tl1, pl, dlm);
    plm = pl;
    pl = pn;
    dlm = dl;
    dl = dn;
    lp1 += 1.0;
    tl1 += 2.0;
  }
  const double k = 1.0 / (4.0 * kPi * r2);
  const double er = k * sa, et = k * sn * sb;
  const double erho = er * sn + et * c;
  ox = rho > 0.0 ? erho * (x / rho) : 0.0;
  oy = rho > 0.0 ? erho * (y / rho) : 0.0;
  oz = er * c - et * sn;
  return true;
}

// One segment's coarse far-rule nodes: Gauss-2 (rule 2), Gauss-4 (1) or
// Gauss-8 (0); the near-axis Taylor branch is tested per node only where the
// segment allows it (axis). nd is the segment's first coarse node.
template <class P>
__device__ __forceinline__ void coarse_segment(double re, double re2, double ze, P nd, int rule, bool axis, RingAcc& acc,
                                               const double2* lt) {
  if (rule == 0) {
    if (axis) {
#pragma unroll 1
      for (int q = 6; q < 14; ++q) {
        const double4 v = load_node(nd, q);
        ring_accum(re, re2, ze - v.y, v.x, v.z, acc, lt);
      }
    } else {
#pragma unroll 2
      for (int q = 6; q < 14; ++q) {
        const double4 v = load_node(nd, q);
        ring_accum_main(re, re2, ze - v.y, v.x, v.w, v.z, acc, lt);
      }
    }
  } else if (rule == 2) {
    const double4 v0 = load_node(nd, 4), v1 = load_node(nd, 5);
    if (axis) {
      ring_accum(re, re2, ze - v0.y, v0.x, v0.z, acc, lt);
      ring_accum(re, re2, ze - v1.y, v1.x, v1.z, acc, lt);
    } else {
      ring_accum_main(re, re2, ze - v0.y, v0.x, v0.w, v0.z, acc, lt);
      ring_accum_main(re, re2, ze - v1.y, v1.x, v1.w, v1.z, acc, lt);
    }
  } else if (axis) {
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
      const double4 v = load_node(nd, q);
      ring_accum(re, re2, ze - v.y, v.x, v.z, acc, lt);
    }
  } else {
#pragma unroll 2
    for (int q = 0; q < 4; ++q) {
      const double4 v = load_node(nd, q);
      ring_accum_main(re, re2, ze - v.y, v.x, v.w, v.z, acc, lt);
    }
  }
}

// Bit c set: the point (re, ze) is at least kappa r_c from cluster c's
// centre, so the cluster's expansion replaces its segments.
__device__ __forceinline__ unsigned long long far_clusters(const BemView& B, double re2, double ze) {
  unsigned long long m = 0ull;
  const cseg_ptr t = (cseg_ptr)B.cl;
  for (int c = 0; c < B.cl_n; ++c) {
    const double dz = ze - t[4 * c];
    if (fma(dz, dz, re2) >= t[4 * c + 2]) m |= 1ull << c;
  }
  return m;
}

// f(j) for the segments j = s, s + S, ... that are not in a far cluster
// (all of them without clusters), in ascending order.
template <class F>
__device__ __forceinline__ void local_segments(const BemView& B, unsigned long long far, int s, int S, F&& f) {
  if (!B.cl) {
    for (int j = s; j < B.n_seg; j += S) f(j);
    return;
  }
  for (int c = 0; c < B.cl_n; ++c) {
    if ((far >> c) & 1ull) continue;
    const int b = B.cl_seg[c], e = B.cl_seg[c + 1];
    for (int j = b + ((s - b) % S + S) % S; j < e; j += S) f(j);
  }
}

// Field of cluster c's exterior expansion at (re, ze), added to the
// accumulator's final-unit terms: phi = (1/4 pi) sum_l m_l q^l P_l(cos t)/r
// about (0, z_c), q = r_c / r, E_r = (1/4 pi r^2) sum m_l (l+1) q^l P_l,
// E_t = (sin t / 4 pi r^2) sum m_l q^l P_l'; terms up to the first l whose
// tail bound sum_{k > l} (k+1) q^k = q^(l+1) (l + 2 - (l+1) q) / (1-q)^2
// is <= cl_tol.
__device__ __forceinline__ void cluster_expansion(const BemView& B, int c, double re, double re2, double ze,
                                                  RingAcc& acc) {
  const cseg_ptr t = (cseg_ptr)B.cl;
  const double dz = ze - t[4 * c];
  const double r2 = fma(dz, dz, re2);
  const double ir = rsqrt64(r2);
  const double cs = dz * ir, sn = re * ir, q = t[4 * c + 1] * ir;
  const double omq = 1.0 - q;
  const double tolw = B.cl_tol * omq * omq;
  const cseg_ptr mom = (cseg_ptr)B.cl_mom + (size_t)c * B.cl_terms;
  const cseg_ptr cf = (cseg_ptr)B.cl_coef;
  double pl = 1.0, plm = 0.0, dl = 0.0, dlm = 0.0, qp = 1.0, sa = 0.0, sb = 0.0;
  double lp1 = 1.0, tl1 = 1.0;  // l + 1 and 2l + 1, exact
  for (int l = 0; l < B.cl_terms; ++l) {
    const double mq = mom[l] * qp;
    sa = fma(mq * lp1, pl, sa);
    sb = fma(mq, dl, sb);
    qp *= q;
    if (qp * fma(-lp1, q, lp1 + 1.0) <= tolw) break;
    const double pn = fma(cf[2 * l] * cs, pl, -cf[2 * l + 1] * plm);
    const double dn = fma(tl1, pl, dlm);
    plm = pl;
    pl = pn;
    dlm = dl;
    dl = dn;
    lp1 += 1.0;
    tl1 += 2.0;
  }
  const double k = (ir * ir) * (1.0 / (4.0 * kPi));
  const double er = k * sa, et = k * sn * sb;
  acc.er_ax += er * sn + et * cs;
  acc.ez_ax += er * cs - et * sn;
}

// Tile holding (rho, z), or -1 (outside the quadtree, or a cell without a
// tile); (cx, cz) and h: the cell's centre and half-size. The halving
// arithmetic is the table builder's (bem_amd.bem.tiles), so both place a
// point in the same cell.
__device__ __forceinline__ int tile_find(const BemView& B, double rho, double z, double& cx, double& cz, double& h) {
  double x0 = B.tl_x0, z0 = B.tl_z0, s = B.tl_size;
  if (!(rho >= x0 && rho < x0 + s && z >= z0 && z < z0 + s)) return -1;
  int k = 0;
  for (int depth = 0; depth <= kTileMaxDepth; ++depth) {
    const int v = B.tl_tree[k];
    if (v < 0) {
      h = 0.5 * s;
      cx = x0 + h;
      cz = z0 + h;
      return v <= -2 ? -2 - v : -1;
    }
    s *= 0.5;
    const int qx = rho >= x0 + s ? 1 : 0, qz = z >= z0 + s ? 1 : 0;
    if (qx) x0 += s;
    if (qz) z0 += s;
    k = v + qx + 2 * qz;
  }
  return -1;
}

// tile_find with a guess: when (rho, z) lies inside tile t's cell by more
// than 1e-10 of its side (far beyond the walk's rounding, so the walk would
// end there too), t and the cell as tile_find gives them, from one load;
// otherwise the walk. t < 0: no guess.
__device__ __forceinline__ int tile_find_hint(const BemView& B, int t, double rho, double z, double& cx, double& cz,
                                              double& h) {
  if (t >= 0 && t < B.tl_n) {
    const double4 c = B.tl_cell[t];
    const double m = 1e-10 * c.z;
    if (rho >= c.x + m && rho < c.x + c.z - m && z >= c.y + m && z < c.y + c.z - m) {
      h = 0.5 * c.z;
      cx = c.x + h;
      cz = c.y + h;
      return t;
    }
  }
  return tile_find(B, rho, z, cx, cz, h);
}

// Rows a = s, s + S, ... of tile t's Chebyshev sums at (u, v) in [-1, 1]^2
// (T_0 = 1, T_1 = x, T_{k+1} = 2x T_k - T_{k-1}): sum_a T_a(u) sum_b
// c[a][b] T_b(v) for E_rho and E_z, added to the accumulator's final-unit
// terms. C: the coefficient pointer type (constant address space where the
// tile is uniform over the wavefront, so the loads are scalar).
template <class C>
__device__ __forceinline__ void tile_rows(C c, double u, double v, int s, int S, RingAcc& acc) {
  double tv[kTileP];
  tv[0] = 1.0;
  tv[1] = v;
#pragma unroll
  for (int b = 2; b < kTileP; ++b) tv[b] = fma(2.0 * v, tv[b - 1], -tv[b - 2]);
  double er = 0.0, ez = 0.0;
  if (S > 1) {
    // Several slices: each runs only its own rows, so a wavefront whose
    // lanes hold different rows waits on ceil(kTileP / S) rows of loads
    // instead of all kTileP. T_a(u) comes from the same recurrence (same
    // values, same order of the sums: the results are unchanged).
    double tu[kTileP];
    tu[0] = 1.0;
    tu[1] = u;
#pragma unroll
    for (int a = 2; a < kTileP; ++a) tu[a] = fma(2.0 * u, tu[a - 1], -tu[a - 2]);
#pragma unroll 1
    for (int a = s; a < kTileP; a += S) {
      double ta = tu[0];
#pragma unroll
      for (int i = 1; i < kTileP; ++i) ta = a == i ? tu[i] : ta;
      double rr = 0.0, rz = 0.0;
#pragma unroll
      for (int b = 0; b < kTileP; ++b) {
        rr = fma(c[a * kTileP + b], tv[b], rr);
        rz = fma(c[kTileP * kTileP + a * kTileP + b], tv[b], rz);
      }
      er = fma(ta, rr, er);
      ez = fma(ta, rz, ez);
    }
    acc.er_ax += er;
    acc.ez_ax += ez;
    return;
  }
  double ta = 1.0, tam = u;  // T_a(u) and T_{a-1}(u) (T_{-1} = T_1)
#pragma unroll 1
  for (int a = 0; a < kTileP; ++a) {
    if (a >= s && (a - s) % S == 0) {
      double rr = 0.0, rz = 0.0;
#pragma unroll
      for (int b = 0; b < kTileP; ++b) {
        rr = fma(c[a * kTileP + b], tv[b], rr);
        rz = fma(c[kTileP * kTileP + a * kTileP + b], tv[b], rz);
      }
      er = fma(ta, rr, er);
      ez = fma(ta, rz, ez);
    }
    const double tn = a == 0 ? u : fma(2.0 * u, ta, -tam);
    tam = ta;
    ta = tn;
  }
  acc.er_ax += er;
  acc.ez_ax += ez;
}

__device__ __forceinline__ void tile_eval(const BemView& B, int t, double u, double v, int s, int S, RingAcc& acc) {
  const int t0 = __builtin_amdgcn_readfirstlane(t);
  if (__ballot(t != t0) == 0ull)
    tile_rows((cseg_ptr)B.tl_coef + (size_t)t0 * kTileCoef, u, v, s, S, acc);
  else
    tile_rows(B.tl_coef + (size_t)t * kTileCoef, u, v, s, S, acc);
}

// Tile of a point whose tile has an empty window (its whole perturbation
// comes from the table), or -1.
__device__ __forceinline__ int tile_only(const BemView& B, double x, double y, double z, double& cx, double& cz,
                                         double& h) {
  if (!B.tl_tree) return -1;
  const int t = tile_find(B, hypot(x, y), z, cx, cz, h);
  if (t < 0) return -1;
  const int2 w = B.tl_win[t];
  return w.x == w.y ? t : -1;
}

// Perturbation at a point of a table-only tile t (tile_only), by one
// thread: every row of the table, then the axis rule of the gather. The
// table comes in by scalar loads where the wavefront's points share it
// (neighbours in the binned slot order do).
__device__ __forceinline__ void tile_point(const BemView& B, int t, double x, double y, double z, double cx, double cz,
                                           double h, double& ox, double& oy, double& oz) {
  const double rho = hypot(x, y), ih = 1.0 / h;
  RingAcc acc;
  acc.zero();
  const double u = (rho - cx) * ih, v = (z - cz) * ih;
  const int t0 = __builtin_amdgcn_readfirstlane(t);
  if (__ballot(t != t0) == 0ull)
    tile_rows((cseg_ptr)B.tl_coef + (size_t)t0 * kTileCoef, u, v, 0, 1, acc);
  else
    tile_rows(B.tl_coef + (size_t)t * kTileCoef, u, v, 0, 1, acc);
  double er, ez;
  ring_finish(acc, rho, er, ez);
  // the azimuth's cosine and sine as x / rho and y / rho (exactly 1 and 0
  // for a point given as (rho, 0, z))
  ox = rho == 0.0 ? 0.0 : er * (x / rho);
  oy = rho == 0.0 ? 0.0 : er * (y / rho);
  oz = ez;
}

__device__ __forceinline__ int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// f(j) for this lane's segments j = s (mod S), ascending, with tiles on: a
// tile lane (tile) its window [jlo, jhi), any other lane every segment
// outside its far clusters (far). The loop runs over the wavefront's hull
// of these ranges and the lanes pick their own segments, so a lane's
// sequence does not depend on its neighbours; with s wavefront-uniform
// (wave-major) so is j, and the table reads in f are scalar loads.
template <class F>
__device__ __forceinline__ void tile_lane_segments(const BemView& B, bool on, bool tile, int jlo, int jhi,
                                                   unsigned long long far, int s, int S, bool uniform, F&& f) {
  int lo = on ? jlo : B.n_seg, hi = on ? jhi : 0;
  lo = wave_min(lo);
  hi = wave_max(hi);
  if (uniform) {
    lo = __builtin_amdgcn_readfirstlane(lo);
    hi = __builtin_amdgcn_readfirstlane(hi);
  }
  for (int j = lo + ((s - lo) % S + S) % S; j < hi; j += S) {
    const bool mine = on && (tile ? (j >= jlo && j < jhi) : !(B.cl && ((far >> B.seg_cl[j]) & 1ull)));
    if (mine) f(j);
  }
}

// Squared distance from (rho, z) to segment g, the projection clamped to the
// segment (reference bem/axisymmetric.py:341-389).
__device__ __forceinline__ double seg_dist2(const double* g, double rho, double z) {
  const double ax = g[S_ABR], ay = g[S_ABZ];
  double t = ((rho - g[S_P0R]) * ax + (z - g[S_P0Z]) * ay) / g[S_ABAB];
  t = fmin(fmax(t, 0.0), 1.0);
  const double dr = rho - (g[S_P0R] + t * ax);
  const double dz = z - (g[S_P0Z] + t * ay);
  return dr * dr + dz * dz;
}

// Bin of a point for the binned lockstep gather: its distance band (d/L in
// [2^(b + kBinBandLo), 2^(b + kBinBandLo + 1)), open-ended at both ends)
// major, its closest segment (>> shift) minor, so consecutive bins are
// neighbouring segments at similar distances.
constexpr int kBinBands = 20;
constexpr int kBinBandLo = -12;
__device__ __forceinline__ int bin_key(int bseg, double bd2, double L, int nsk, int shift) {
  const double r = bd2 / (L * L);
  int band = 0;
  if (!(r < 1e300)) {
    band = kBinBands - 1;
  } else if (r > 0.0) {
    int e;
    (void)frexp(r, &e);              // log2 r in [e - 1, e)
    band = ((e - 1) >> 1) - kBinBandLo;  // floor(log2 sqrt r) - kBinBandLo
    band = band < 0 ? 0 : (band >= kBinBands ? kBinBands - 1 : band);
  }
  return band * nsk + (bseg >> shift);
}

// Perturbation field -grad S[sigma] at m <= NMAX points held in LDS, by all
// T threads of the block: S threads ("slices") per point, slice s taking
// segments s, s+S, ... Two lane layouts: slice-major (S = T/m, point p's
// slices are threads p*S .. p*S+S-1) and, for 32 < m <= 64, wave-major
// (lane = point, wavefront = slice, S = T/64): there every wavefront walks
// one segment sequence, so its node and segment reads are uniform scalar
// loads instead of 64 scattered vector reads per node. Closest segment
// first (first index wins ties, reference bem/axisymmetric.py:341-389) to
// place exactly-on-surface points 1e-4 L off the surface (554-557), then
// per segment either its far Gauss nodes or its near sinh rule, the latter
// shared out over the point's slices afterwards. Slice partials are
// combined by a fixed-pairing tree, so results are deterministic (they
// depend on S, i.e. on the launch shape, in the last bits). m must be
// block-uniform. Far-rule gathers pick each point's rule per lane, so a
// point's result does not depend on the other points of its block. BIN:
// the binned lockstep gather's form, with the points valid[p] != 0 of
// 0..NMAX-1 and each point's bin (bin_key, or key_tile0 + its tile)
// written to okey[p]: for NMAX = 64 wave-major (lane = point, so the result
// is independent of m too), for NMAX = 16 slice-major with S = T/16 slices
// per point whatever the number of valid points (the few-records form:
// lower latency per point). With tiles on (far-rule gathers), a point inside a tile searches
// and sums only its tile's window and adds the tile's table: the window
// holds every segment within max(alpha s, L) of the cell, so the closest
// segment wherever the on-surface offset applies, and every segment the
// near rule would take, are among them.
template <int T, int NMAX, int PR, bool FAR, bool BIN = false>
__device__ void bem_pert_block(const BemView& B, int m, const double* px, const double* py, const double* pz,
                               double* ox, double* oy, double* oz, BemShared<T, NMAX>& sh, const int* valid = nullptr,
                               int* okey = nullptr, int key_nsk = 0, int key_shift = 0, int key_tile0 = 0,
                               const int* ptile = nullptr) {
  static_assert(T % 64 == 0, "whole wavefronts");
  static_assert(!BIN || (FAR && (NMAX == 64 || NMAX == 16)), "the binned gather is a far-rule gather of 64 or 16 points");
  // The few-records form's slice reductions by shuffles: only where a
  // point's T / 16 slices are aligned lanes of one wavefront (not T = 192,
  // whose 12 slices per point straddle wavefronts).
  constexpr bool kFewShfl = BIN && NMAX == 16 && 64 % (T / 16) == 0;
  if (m <= 0) return;
  fill_log_table<T>(sh.logt);  // read after the barrier that ends the closest-segment pass
  const int tid = threadIdx.x;
  const bool wave_major = BIN ? NMAX == 64 : (m > 32 && m <= 64 && B.n_seg >= 4 * (T / 64));
  const int S = wave_major ? T / 64 : (BIN ? T / NMAX : (m >= T ? 1 : T / m));
  const int p = wave_major ? (tid & 63) : tid / S;
  const int s = wave_major ? (tid >> 6) : tid - p * S;
  const bool on = p < m && (!BIN || valid[p] != 0);
  // thread holding slice k of this thread's point: slot0 + k * sstep
  const int slot0 = wave_major ? p : p * S, sstep = wave_major ? 64 : 1;
  auto slot = [=](int k) { return slot0 + k * sstep; };
  const double nw = FAR ? B.near_w : kRefNearWidth;  // near-rule interval width
  // Tiles: this thread's point's tile (-1: none), its window and cell.
  const bool tiles = FAR && B.tl_tree != nullptr;
  const int su = wave_major ? __builtin_amdgcn_readfirstlane(s) : s;
  int tleaf = -1, jlo = 0, jhi = B.n_seg;
  double tcx = 0.0, tcz = 0.0, th = 1.0;
  if (tiles && on) {
    if (BIN && ptile && ptile[p] < B.tl_n) {
      // the tile the packing found for this very position (tile_find's
      // result; the cell as the walk leaves it)
      tleaf = ptile[p];
      if (tleaf >= 0) {
        const double4 c = B.tl_cell[tleaf];
        th = 0.5 * c.z;
        tcx = c.x + th;
        tcz = c.y + th;
      }
    } else {
      tleaf = tile_find(B, hypot(px[p], py[p]), pz[p], tcx, tcz, th);
    }
    if (tleaf >= 0) {
      const int2 w = B.tl_win[tleaf];
      jlo = w.x;
      jhi = w.y;
    }
  }
  const bool tl = tleaf >= 0;
  if (tiles) {
    // all lanes together (tile_lane_segments reduces over the wavefront)
    double rho = 0.0, z = 0.0;
    if (on) {
      rho = BIN ? px[p] : hypot(px[p], py[p]);  // the binned gather's points are (rho, 0, z)
      z = pz[p];
    }
    double bd2 = INFINITY;
    int bseg = 0;
    const unsigned long long farc = (on && !tl && B.cl) ? far_clusters(B, rho * rho, z) : 0ull;
    tile_lane_segments(B, on, tl, jlo, jhi, farc, su, S, wave_major, [&](int j) {
      const double d2 = seg_dist2(B.segs + (size_t)j * SEG_STRIDE, rho, z);
      if (d2 < bd2) {
        bd2 = d2;
        bseg = j;
      }
    });
    if (kFewShfl) {
      // few-records form: a point's slices are aligned lanes of one
      // wavefront, so the minimum is taken by shuffles, without barriers
      for (int h = 1; h < S; h <<= 1) {
        const double d = __shfl_xor(bd2, h);
        const int sg = __shfl_xor(bseg, h);
        if (d < bd2 || (d == bd2 && sg < bseg)) {
          bd2 = d;
          bseg = sg;
        }
      }
    }
    if (on) {
      sh.u.a.bd2[tid] = bd2;
      sh.u.a.bseg[tid] = bseg;
    }
  } else if (on) {
    const double x = px[p], y = py[p], z = pz[p];
    const double rho = hypot(x, y);
    double bd2 = INFINITY;
    int bseg = 0;
    auto closest = [&](int j) {
      const double d2 = seg_dist2(B.segs + (size_t)j * SEG_STRIDE, rho, z);
      if (d2 < bd2) {
        bd2 = d2;
        bseg = j;
      }
    };
    if (FAR) {
      // the closest segment of a point on the surface lies in a cluster
      // whose sphere holds the point, never in a far one: only the
      // offset of such points depends on this search
      local_segments(B, B.cl ? far_clusters(B, rho * rho, z) : 0ull, s, S, closest);
    } else {
      for (int j = s; j < B.n_seg; j += S) closest(j);
    }
    sh.u.a.bd2[tid] = bd2;
    sh.u.a.bseg[tid] = bseg;
  }
  // Closest segment over the point's slices: a fixed-pairing tree (slot(s)
  // is this thread; the (distance, index) minimum does not depend on the
  // pairing). The few-records form with tiles has it already (shuffles);
  // there slice 0 reads back only its own LDS slot below.
  const bool shfl = kFewShfl && tiles;
  for (int h = 1; h < S && !shfl; h <<= 1) {
    __syncthreads();
    if (on && (s & (2 * h - 1)) == 0 && s + h < S) {
      const double d = sh.u.a.bd2[slot(s + h)];
      const int sg = sh.u.a.bseg[slot(s + h)];
      if (d < sh.u.a.bd2[tid] || (d == sh.u.a.bd2[tid] && sg < sh.u.a.bseg[tid])) {
        sh.u.a.bd2[tid] = d;
        sh.u.a.bseg[tid] = sg;
      }
    }
  }
  if (!shfl) __syncthreads();
  if (on && s == 0) {
    const double bd2 = sh.u.a.bd2[tid];
    const int bseg = sh.u.a.bseg[tid];
    const double x = px[p], y = py[p], z = pz[p];
    const double rho = BIN ? x : hypot(x, y);
    const double* g = B.segs + (size_t)bseg * SEG_STRIDE;
    const double L = g[S_L];
    const double thr = 1e-9 * L;
    const double off = bd2 < thr * thr ? 1e-4 * L : 0.0;
    if (BIN) okey[p] = tl ? key_tile0 + tleaf : bin_key(bseg, bd2, L, key_nsk, key_shift);
    double re = rho + off * g[S_NR];
    if (rho == 0.0) re = 0.0;
    sh.re[p] = re;
    sh.ze[p] = z + off * g[S_NZ];
    sh.rho[p] = rho;
    if (FAR) {  // x / rho and y / rho: exactly 1 and 0 for the binned gather's (rho, 0, z) points
      sh.cph[p] = rho == 0.0 ? 1.0 : x / rho;
      sh.sph[p] = rho == 0.0 ? 0.0 : y / rho;
    } else {
      const double phi = atan2(y, x);
      sh.cph[p] = cos(phi);
      sh.sph[p] = sin(phi);
    }
  }
  __syncthreads();
  RingAcc acc;
  acc.zero();
  RingAcc32 acc32;
  acc32.zero();
  const double re = on ? sh.re[p] : 0.0, ze = on ? sh.ze[p] : 0.0, re2 = re * re;
  int n_near = 0;  // near segments handed to the point's cooperative pass
  const unsigned long long farm = (FAR && B.cl && on && !tl) ? far_clusters(B, re2, ze) : 0ull;
  // One segment's contribution to this thread's point (its far nodes, or
  // its near rule handed to the cooperative pass).
  auto seg = [&](int j) {
        const cseg_ptr g = (cseg_ptr)B.segs + (size_t)j * SEG_STRIDE;
        const double L = g[S_L];
        double proj, dist;
        segment_proj(re, ze, g, proj, dist);
        if (FAR) {
          // Far rule, picked per lane: lanes needing different rules run
          // them one after the other (each a branch with scalar table
          // loads), so a point's sum does not depend on its neighbours.
          // The near-axis test is per wavefront: it changes no result.
          const bool near = dist < L;
          if (near) {
            if (n_near < kNearSlots)
              sh.u.n.seg[n_near++][tid] = j;
            else
              acc.add(near_segment(re, re2, ze, B.segs + (size_t)j * SEG_STRIDE, B.sigma[j], B.sigma[j + 1], proj,
                                   dist, sh.logt, nw));
          }
          const int rl = near ? 3 : (dist >= B.k2 * L ? 2 : (dist >= B.k4 * L ? 1 : 0));
          const bool axw =
              __ballot(!near && ring_axis_possible(re2, ze, g[S_P0R], g[S_P0Z], g[S_ABR], g[S_ABZ])) != 0ull;
          const cnode_ptr cn = (cnode_ptr)B.coarse + (size_t)j * kCoarsePer;
          if (rl == 0)
            coarse_segment(re, re2, ze, cn, 0, axw, acc, sh.logt);
          else if (rl == 1)
            coarse_segment(re, re2, ze, cn, 1, axw, acc, sh.logt);
          else if (rl == 2)
            coarse_segment(re, re2, ze, cn, 2, axw, acc, sh.logt);
        } else if (dist < L) {
          if (n_near < kNearSlots)
            sh.u.n.seg[n_near++][tid] = j;
          else
            acc.add(near_segment(re, re2, ze, B.segs + (size_t)j * SEG_STRIDE, B.sigma[j], B.sigma[j + 1], proj,
                                 dist, sh.logt, nw));
        } else if (PR == SPT_PREC_MIXED && dist >= 2.0 * L) {
          const cnode_ptr nd = (cnode_ptr)B.nodes + (size_t)j * B.nq;
#pragma unroll 1
          for (int q = 0; q < B.nq; ++q) {
            const double4 v = load_node(nd, q);
            ring_accum32(re, re2, ze - v.y, v.x, (float)v.z, acc32, acc);
          }
        } else {
          // Segments that cannot meet the near-axis branch take the
          // branch-free main term, two nodes per trip (nq is a multiple of 4).
          const cnode_ptr nd = (cnode_ptr)B.nodes + (size_t)j * B.nq;
          if (ring_axis_possible(re2, ze, g[S_P0R], g[S_P0Z], g[S_ABR], g[S_ABZ])) {
#pragma unroll 1
            for (int q = 0; q < B.nq; ++q) {
              const double4 v = load_node(nd, q);
              ring_accum(re, re2, ze - v.y, v.x, v.z, acc, sh.logt);
            }
          } else {
#pragma unroll 2
            for (int q = 0; q < B.nq; ++q) {
              const double4 v = load_node(nd, q);
              ring_accum_main(re, re2, ze - v.y, v.x, v.w, v.z, acc, sh.logt);
            }
          }
        }
  };
  if (tiles && B.n_seg >= 4 * S && BIN && NMAX == 16) {
    // Few-records form. Lanes outside a tile: every segment outside their
    // far clusters. Tile lanes: the far nodes of the window's segments are
    // numbered consecutively (window order, then node order) and node u
    // goes to slice u mod S, so each slice runs about one node per segment
    // instead of one slice running a segment's whole rule while the
    // point's other slices wait (a lone block's latency is this chain); a
    // near segment j goes to the cooperative pass from slice j mod S, as
    // below.
    tile_lane_segments(B, on && !tl, false, 0, B.n_seg, farm, su, S, false, seg);
    const int wn = (on && tl) ? jhi - jlo : 0;
    const int wmax = wave_max(wn);
    int u0 = 0;
    for (int k = 0; k < wmax; ++k) {
      if (k < wn) {
        const int j = jlo + k;
        const double* g = B.segs + (size_t)j * SEG_STRIDE;
        const double L = g[S_L];
        double proj, dist;
        segment_proj(re, ze, g, proj, dist);
        if (dist < L) {
          if (j % S == s) {
            if (n_near < kNearSlots)
              sh.u.n.seg[n_near++][tid] = j;
            else
              acc.add(near_segment(re, re2, ze, g, B.sigma[j], B.sigma[j + 1], proj, dist, sh.logt, nw));
          }
        } else {
          const int rl = dist >= B.k2 * L ? 2 : (dist >= B.k4 * L ? 1 : 0);
          const int nq = 8 >> rl, q0 = rl == 0 ? 6 : (rl == 1 ? 0 : 4);  // Gauss-8, -4, -2 (coarse_segment)
          const bool axw =
              __ballot(ring_axis_possible(re2, ze, g[S_P0R], g[S_P0Z], g[S_ABR], g[S_ABZ])) != 0ull;
          const double4* cn = B.coarse + (size_t)j * kCoarsePer + q0;
          for (int q = ((s - u0) % S + S) % S; q < nq; q += S) {
            const double4 v = cn[q];
            if (axw)
              ring_accum(re, re2, ze - v.y, v.x, v.z, acc, sh.logt);
            else
              ring_accum_main(re, re2, ze - v.y, v.x, v.w, v.z, acc, sh.logt);
          }
          u0 += nq;
        }
      }
    }
  } else if (tiles && B.n_seg >= 4 * S) {
    // tile lanes: their window; the others: outside their far clusters
    tile_lane_segments(B, on, tl, jlo, jhi, farm, su, S, wave_major, seg);
  } else if (on) {
    // Many segments per slice: stride over whole segments from j0. In the
    // wave-major layout j0 is made wavefront-uniform (readfirstlane), so
    // the table reads in seg compile to scalar loads. Far-rule gathers skip
    // the segments of far clusters (their expansions are added below).
    auto whole_segments = [&](int j0) {
      if (FAR)
        local_segments(B, farm, j0, S, seg);
      else
        for (int j = j0; j < B.n_seg; j += S) seg(j);
    };
    if (B.n_seg >= 4 * S) {
      if (wave_major)
        whole_segments(__builtin_amdgcn_readfirstlane(s));
      else
        whole_segments(s);
    } else {
      // Few segments per slice: stride over sub-panels (4 Gauss nodes) so
      // slices stay balanced. The near test is per segment; the unit with
      // sub == 0 of a near segment hands the segment to the cooperative
      // near pass, the segment's other units skip.
      const int n_sub = B.nq >> 2;
      const int n_units = B.n_seg * n_sub;
      for (int u = s; u < n_units; u += S) {
        const int j = u / n_sub;
        const int sub = u - j * n_sub;
        if (FAR && B.cl && ((farm >> B.seg_cl[j]) & 1ull)) continue;
        if (tl && (j < jlo || j >= jhi)) continue;  // outside the tile's window: in its table
        const double* g = B.segs + (size_t)j * SEG_STRIDE;
        const double L = g[S_L];
        double proj, dist;
        segment_proj(re, ze, g, proj, dist);
        if (dist < L) {
          if (sub == 0) {
            if (n_near < kNearSlots)
              sh.u.n.seg[n_near++][tid] = j;
            else
              acc.add(near_segment(re, re2, ze, g, B.sigma[j], B.sigma[j + 1], proj, dist, sh.logt, nw));
          }
        } else if (FAR) {
          if (sub == 0)
            coarse_segment(re, re2, ze, B.coarse + (size_t)j * kCoarsePer,
                           dist >= B.k2 * L ? 2 : (dist >= B.k4 * L ? 1 : 0),
                           ring_axis_possible(re2, ze, g[S_P0R], g[S_P0Z], g[S_ABR], g[S_ABZ]), acc, sh.logt);
        } else if (PR == SPT_PREC_MIXED && dist >= 2.0 * L) {
          const double4* nd = B.nodes + (size_t)u * 4;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double4 v = nd[q];
            ring_accum32(re, re2, ze - v.y, v.x, (float)v.z, acc32, acc);
          }
        } else {
          const double4* nd = B.nodes + (size_t)u * 4;
          if (ring_axis_possible(re2, ze, g[S_P0R], g[S_P0Z], g[S_ABR], g[S_ABZ])) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const double4 v = nd[q];
              ring_accum(re, re2, ze - v.y, v.x, v.z, acc, sh.logt);
            }
          } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const double4 v = nd[q];
              ring_accum_main(re, re2, ze - v.y, v.x, v.x * v.x, v.z, acc, sh.logt);  // a^2 formed here: 24-byte loads
            }
          }
        }
      }
    }
  }
  if (FAR && B.cl && on) {
    for (int c = s; c < B.cl_n; c += S)
      if ((farm >> c) & 1ull) cluster_expansion(B, c, re, re2, ze, acc);
  }
  if (tl) {
    // the tile's table at the (possibly offset) evaluation point
    const double ih = 1.0 / th;
    tile_eval(B, tleaf, (re - tcx) * ih, (ze - tcz) * ih, su, S, acc);
  }
  sh.u.n.cnt[tid] = n_near;
  {
    const unsigned long long lanes = __ballot(n_near > 0);
    if ((tid & 63) == 0) sh.nmask[tid >> 6] = lanes;
  }
  __syncthreads();
  if (on) {
    // Cooperative near rule: the Gauss nodes of all the point's handed-over
    // near segments (slice order, then list order), numbered consecutively,
    // node f to slice f mod S. Balances the sinh rule, which would
    // otherwise fall on the one or two slices holding the point's closest
    // segments.
    int base = 0;
    // far-rule gathers with 2 or 4 slices per point take the pair form:
    // every slice then holds the same pairs of each near segment
    const bool pairs = FAR && (S == 2 || S == 4);
    auto near_slice = [&](int t) {
      const int c = sh.u.n.cnt[t];
      for (int k = 0; k < c; ++k) {
        const int j = sh.u.n.seg[k][t];
        const double* g = B.segs + (size_t)j * SEG_STRIDE;
        double proj, dist;
        segment_proj(re, ze, g, proj, dist);
        const NearRule r = FAR ? near_rule_fast(g, proj, dist, nw, sh.logt) : near_rule(g, proj, dist, nw);
        if (pairs) {
          near_pairs(re, re2, ze, g, B.sigma[j], B.sigma[j + 1], r, su, S, acc, sh.logt);
        } else {
          near_nodes(re, re2, ze, g, B.sigma[j], B.sigma[j + 1], r, (s - base % S + S) % S, S, acc, sh.logt);
          base += 8 * r.count;
        }
      }
    };
    if (wave_major) {
      for (int ss = 0; ss < S; ++ss) near_slice(slot(ss));
    } else {
      // The point's slices are threads slot0 .. slot0 + S - 1: visit only
      // those whose wavefront ballot bit is set, in thread (= slice) order.
      const int t_end = slot0 + S;
      for (int w = slot0 >> 6; w <= (t_end - 1) >> 6; ++w) {
        unsigned long long lanes = sh.nmask[w];
        const int lo = slot0 - 64 * w, hi = t_end - 64 * w;
        if (lo > 0) lanes &= ~0ull << lo;
        if (hi < 64) lanes &= (1ull << hi) - 1ull;
        while (lanes) {
          const int b = __builtin_ctzll(lanes);
          lanes &= lanes - 1ull;
          near_slice(64 * w + b);
        }
      }
    }
    if (PR == SPT_PREC_MIXED) {
      acc.er_main += acc32.er.value();
      acc.ez_main += acc32.ez.value();
    }
  }
  if (kFewShfl) {
    // few-records form: the slices are aligned lanes of one wavefront;
    // the tree below by shuffles (lane s takes s ^ h, i.e. s + h where the
    // tree adds: the same pairing and order, so the same sums)
    double v[4] = {acc.er_main, acc.ez_main, acc.er_ax, acc.ez_ax};
    for (int h = 1; h < S; h <<= 1) {
#pragma unroll
      for (int c = 0; c < 4; ++c) v[c] += __shfl_xor(v[c], h);
    }
    if (on && s == 0) {
      const RingAcc tot = {v[0], v[1], v[2], v[3]};
      double er, ez;
      ring_finish(tot, sh.re[p], er, ez);
      const bool axis = sh.rho[p] == 0.0;
      ox[p] = axis ? 0.0 : er * sh.cph[p];
      oy[p] = axis ? 0.0 : er * sh.sph[p];
      oz[p] = ez;
    }
    __syncthreads();
    return;
  }
  __syncthreads();  // near-list reads done before acc (aliased) is written
  sh.u.acc[0][tid] = acc.er_main;
  sh.u.acc[1][tid] = acc.ez_main;
  sh.u.acc[2][tid] = acc.er_ax;
  sh.u.acc[3][tid] = acc.ez_ax;
  // Slice partials summed by a fixed-pairing tree: deterministic.
  for (int h = 1; h < S; h <<= 1) {
    __syncthreads();
    if (on && (s & (2 * h - 1)) == 0 && s + h < S) {
#pragma unroll
      for (int c = 0; c < 4; ++c) sh.u.acc[c][tid] += sh.u.acc[c][slot(s + h)];
    }
  }
  __syncthreads();
  if (on && s == 0) {
    const RingAcc tot = {sh.u.acc[0][tid], sh.u.acc[1][tid], sh.u.acc[2][tid], sh.u.acc[3][tid]};
    double er, ez;
    ring_finish(tot, sh.re[p], er, ez);
    const bool axis = sh.rho[p] == 0.0;
    ox[p] = axis ? 0.0 : er * sh.cph[p];
    oy[p] = axis ? 0.0 : er * sh.sph[p];
    oz[p] = ez;
  }
  __syncthreads();
}

// The gather with the reference's fixed rule, or with the far rule when the
// view carries its coarse table (two instantiations, so the fixed rule's
// code is unchanged by the far rule's).
template <int T, int NMAX, int PR>
__device__ __forceinline__ void bem_pert_any(const BemView& B, int m, const double* px, const double* py,
                                             const double* pz, double* ox, double* oy, double* oz,
                                             BemShared<T, NMAX>& sh) {
  if (B.coarse)
    bem_pert_block<T, NMAX, PR, true>(B, m, px, py, pz, ox, oy, oz, sh);
  else
    bem_pert_block<T, NMAX, PR, false>(B, m, px, py, pz, ox, oy, oz, sh);
}

// ------------------------------------------------------ geometry helpers
// Smoothstep weight: 1 at z <= z_lo, 0 at z >= z_hi (reference
// bem/image_charge.py:328-342, used by bem/geometry.py:44-55).
__device__ __forceinline__ double blend_weight(double z, double z_lo, double z_hi) {
  double t = (z_hi - z) / (z_hi - z_lo);
  t = fmin(fmax(t, 0.0), 1.0);
  return t * t * (3.0 - 2.0 * t);
}

// |p - nearest point of the axis segment [0, H]|^2: the capsule-interior
// measure of the post (H = 0: squared distance from the origin, as the
// reference hemisphere's sum(position**2)).
__device__ __forceinline__ double post_r2(double x, double y, double z, double H) {
  if (H == 0.0) return x * x + y * y + z * z;
  const double zc = fmin(fmax(z, -H), H);
  const double dz = z - zc;
  return x * x + y * y + dz * dz;
}

__host__ __device__ inline bool is_image_bem(int image) {
  return image == SPT_IMAGE_BEM || image == SPT_IMAGE_BEM_MIRROR;
}

__device__ __forceinline__ bool is_bem(const spt_geometry& G) {
  return G.kind == SPT_GEOM_BEM_POST || G.kind == SPT_GEOM_BEM_WELL;
}

// Inside (or on) the real cylindrical-well conductor (reference
// bem/fields.py:89-135): below the plane outside the well radius, below
// the bottom inside it; within a fillet's square the distance to the
// fillet centre decides, biased 1e-6 R towards inside.
__device__ __forceinline__ bool well_contains(const spt_geometry& G, double rho, double z) {
  const double R = G.R, H = G.H;
  const double eps = 1e-6 * R;
  bool m = rho >= R * (1.0 - 1e-6) ? (z <= 0.0) : (z <= -H * (1.0 + 1e-6));
  const double rf = G.rim_fillet, bf = G.bottom_fillet;
  if (rf > 0.0 && rho >= R - eps && rho <= R + rf + eps && z >= -rf - eps && z <= eps) {
    const double dr = rho - (R + rf), dz = z + rf;
    m = sqrt(dr * dr + dz * dz) <= rf;
  }
  if (bf > 0.0 && rho <= R + eps && rho >= R - bf - eps && z <= -H + bf + eps && z >= -H - eps) {
    const double dr = rho - (R - bf), dz = z - (-H + bf);
    m = sqrt(dr * dr + dz * dz) >= bf;
  }
  return m;
}

// Outside the conductor, where the BEM field is defined (reference
// bem/fields.py:200-211 tip, 338-352 well; the post measures r from its
// axis segment).
__device__ __forceinline__ bool bem_valid(const spt_geometry& G, double x, double y, double z) {
  if (G.kind == SPT_GEOM_BEM_WELL) return !well_contains(G, sqrt(x * x + y * y), z);
  const double r = sqrt(post_r2(x, y, z, G.H));
  return (r >= G.R * (1.0 - 1e-6)) && (z >= 0.0);
}

// Does this point need the BEM perturbation? (valid region and w > 0)
__device__ __forceinline__ bool bem_needed(const spt_geometry& G, double x, double y, double z) {
  return bem_valid(G, x, y, z) && blend_weight(z, 0.5 * G.z_handoff, G.z_handoff) > 0.0;
}

// Full external field at one point; (qx, qy, qz) is the BEM perturbation
// when bem_needed() held, else unused.
__device__ void geometry_field_point(const spt_geometry& G, double x, double y, double z, double qx, double qy,
                                     double qz, double& Ex, double& Ey, double& Ez) {
  if (G.kind == SPT_GEOM_FLAT) {  // GunField (reference fields.py:43-59)
    Ex = 0.0;
    Ey = 0.0;
    Ez = z >= 0.0 ? G.E_gun : 0.0;
    return;
  }
  if (G.kind == SPT_GEOM_TIP) {  // reference fields.py:119-160
    const double R = G.R;
    const double r = sqrt(x * x + y * y + z * z);
    const bool valid = (r >= R * (1.0 - 1e-6)) && (z >= 0.0);
    const double rs = valid ? r : R;
    const double r3 = rs * rs * rs;
    const double r5 = r3 * rs * rs;
    const double R3 = R * R * R;
    if (!valid) {
      Ex = Ey = Ez = 0.0;
      return;
    }
    Ex = 3.0 * G.E_gun * R3 * x * z / r5;
    Ey = 3.0 * G.E_gun * R3 * y * z / r5;
    Ez = G.E_gun * (1.0 - R3 / r3 + 3.0 * R3 * z * z / r5);
    return;
  }
  // BEM post / hemisphere / well: reference bem/fields.py:200-211,338-352,
  // bem/geometry.py:319-324,779-784
  const bool valid = bem_valid(G, x, y, z);
  const double w = blend_weight(z, 0.5 * G.z_handoff, G.z_handoff);
  double nx = 0.0, ny = 0.0, nz = 0.0;
  if (valid && w > 0.0) {
    nx = 0.0 + qx;
    ny = 0.0 + qy;
    nz = G.E_gun + qz;
  }
  const double fz = z >= 0.0 ? G.E_gun : 0.0;
  Ex = w * nx + (1.0 - w) * 0.0;
  Ey = w * ny + (1.0 - w) * 0.0;
  Ez = w * nz + (1.0 - w) * fz;
}

// Geometry.kill_mask, optionally OR z >= z_max (reference geometry.py:152-155,
// 238-248,304-322; bem/geometry.py:429-436).
__device__ __forceinline__ bool kill_point(const spt_geometry& G, double x, double y, double z) {
  bool k = false;
  if (G.kind == SPT_GEOM_BEM_WELL) {
    k = well_contains(G, sqrt(x * x + y * y), z);
  } else if (G.kind != SPT_GEOM_FLAT) {
    const double rk = G.R * (1.0 - 1e-6);
    k = post_r2(x, y, z, G.kind == SPT_GEOM_BEM_POST ? G.H : 0.0) <= rk * rk;
  }
  if (G.has_kill_z) k = k || (z <= G.kill_z_below);
  if (G.has_z_max) k = k || (z >= G.z_max);
  return k;
}

// One Plummer-softened pair term (reference forces.py:16-29).
__device__ __forceinline__ void pair_term(double xi, double yi, double zi, double qi, double xj, double yj, double zj,
                                          double qj, double a2, double& fx, double& fy, double& fz) {
  const double dx = xi - xj, dy = yi - yj, dz = zi - zj;
  const double r2 = dx * dx + dy * dy + dz * dz + a2;
  const double r3 = r2 * sqrt(r2);
  const double c = kCoulomb * qi * qj / r3;
  fx += c * dx;
  fy += c * dy;
  fz += c * dz;
}

// FP32 pair term with Kahan accumulation: separations in FP64, then
// rounded; k q_i q_j rounded once.
__device__ __forceinline__ void pair_term32(double xi, double yi, double zi, double qi, double xj, double yj,
                                            double zj, double qj, double a2, Kahan32& fx, Kahan32& fy,
                                            Kahan32& fz) {
  const float dx = (float)(xi - xj), dy = (float)(yi - yj), dz = (float)(zi - zj);
  const float r2 = dx * dx + dy * dy + dz * dz + (float)a2;
  const float r3 = r2 * sqrtf(r2);
  const float c = (float)(kCoulomb * qi * qj) / r3;
  fx.add(c * dx);
  fy.add(c * dy);
  fz.add(c * dz);
}

// In-group Coulomb and image forces on particle i from LDS arrays, FP32
// with compensated sums (mixed precision).
__device__ void pair_forces_i32(const spt_geometry& G, int i, int n, const double* sx, const double* sy,
                                const double* sz, const double* sq, const int* act, double a2, double* Fc,
                                double* Fi) {
  Fc[0] = Fc[1] = Fc[2] = 0.0;
  Fi[0] = Fi[1] = Fi[2] = 0.0;
  if (!act[i]) return;
  const double xi = sx[i], yi = sy[i], zi = sz[i], qi = sq[i];
  Kahan32 cx, cy, cz;
  cx.zero();
  cy.zero();
  cz.zero();
  for (int j = 0; j < n; ++j) {
    if (!act[j]) continue;
    pair_term32(xi, yi, zi, qi, sx[j], sy[j], sz[j], sq[j], a2, cx, cy, cz);
  }
  Fc[0] = cx.value();
  Fc[1] = cy.value();
  Fc[2] = cz.value();
  if (G.image != SPT_IMAGE_PLANE && G.image != SPT_IMAGE_TIP3) return;  // the BEM image models have their own kernels
  Kahan32 ix, iy, iz;
  ix.zero();
  iy.zero();
  iz.zero();
  if (G.image == SPT_IMAGE_PLANE) {
    for (int j = 0; j < n; ++j) {
      if (!act[j]) continue;
      pair_term32(xi, yi, zi, qi, sx[j], sy[j], -2.0 * G.z0 - sz[j], -sq[j], a2, ix, iy, iz);
    }
  } else {
    const double a = G.R - G.z0;
    for (int j = 0; j < n; ++j) {
      if (!act[j]) continue;
      const double xj = sx[j], yj = sy[j], zj = sz[j], qj = sq[j];
      double d = sqrt(xj * xj + yj * yj + zj * zj);
      d = d > 0.0 ? d : a;
      const double sc = (a * a) / (d * d);
      const double six = xj * sc, siy = yj * sc, siz = zj * sc;
      const double siq = -qj * (a / d);
      pair_term32(xi, yi, zi, qi, six, siy, siz, siq, a2, ix, iy, iz);
      pair_term32(xi, yi, zi, qi, xj, yj, -2.0 * G.z0 - zj, -qj, a2, ix, iy, iz);
      pair_term32(xi, yi, zi, qi, six, siy, -2.0 * G.z0 - siz, -siq, a2, ix, iy, iz);
    }
  }
  Fi[0] = ix.value();
  Fi[1] = iy.value();
  Fi[2] = iz.value();
}

template <int PR>
__device__ __forceinline__ void pair_forces_p(const spt_geometry& G, int i, int n, const double* sx,
                                              const double* sy, const double* sz, const double* sq, const int* act,
                                              double a2, double* Fc, double* Fi);

// In-group Coulomb and image forces on particle i from LDS arrays
// (reference forces.py:32-191). Inactive partners contribute nothing.
__device__ void pair_forces_i(const spt_geometry& G, int i, int n, const double* sx, const double* sy,
                              const double* sz, const double* sq, const int* act, double a2, double* Fc, double* Fi) {
  Fc[0] = Fc[1] = Fc[2] = 0.0;
  Fi[0] = Fi[1] = Fi[2] = 0.0;
  if (!act[i]) return;
  const double xi = sx[i], yi = sy[i], zi = sz[i], qi = sq[i];
  for (int j = 0; j < n; ++j) {
    if (!act[j]) continue;
    pair_term(xi, yi, zi, qi, sx[j], sy[j], sz[j], sq[j], a2, Fc[0], Fc[1], Fc[2]);
  }
  if (G.image == SPT_IMAGE_PLANE) {
    for (int j = 0; j < n; ++j) {
      if (!act[j]) continue;
      pair_term(xi, yi, zi, qi, sx[j], sy[j], -2.0 * G.z0 - sz[j], -sq[j], a2, Fi[0], Fi[1], Fi[2]);
    }
  } else if (G.image == SPT_IMAGE_TIP3) {
    const double a = G.R - G.z0;
    double f1[3] = {0, 0, 0}, f2[3] = {0, 0, 0}, f3[3] = {0, 0, 0};
    for (int j = 0; j < n; ++j) {
      if (!act[j]) continue;
      const double xj = sx[j], yj = sy[j], zj = sz[j], qj = sq[j];
      double d = sqrt(xj * xj + yj * yj + zj * zj);
      d = d > 0.0 ? d : a;
      const double sc = (a * a) / (d * d);
      const double six = xj * sc, siy = yj * sc, siz = zj * sc;
      const double siq = -qj * (a / d);
      pair_term(xi, yi, zi, qi, six, siy, siz, siq, a2, f1[0], f1[1], f1[2]);
      pair_term(xi, yi, zi, qi, xj, yj, -2.0 * G.z0 - zj, -qj, a2, f2[0], f2[1], f2[2]);
      pair_term(xi, yi, zi, qi, six, siy, -2.0 * G.z0 - siz, -siq, a2, f3[0], f3[1], f3[2]);
    }
    for (int c = 0; c < 3; ++c) Fi[c] = f1[c] + f2[c] + f3[c];
  }
}

template <int PR>
__device__ __forceinline__ void pair_forces_p(const spt_geometry& G, int i, int n, const double* sx,
                                              const double* sy, const double* sz, const double* sq, const int* act,
                                              double a2, double* Fc, double* Fi) {
  if (PR == SPT_PREC_MIXED)
    pair_forces_i32(G, i, n, sx, sy, sz, sq, act, a2, Fc, Fi);
  else
    pair_forces_i(G, i, n, sx, sy, sz, sq, act, a2, Fc, Fi);
}

// ----------------------------------------------------------- screens
__device__ __forceinline__ double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : (x == 0.0 ? 0.0 : x)); }

__device__ __forceinline__ void hbasis(double s, double& h00, double& h10, double& h01, double& h11) {
  const double s2 = s * s, s3 = s2 * s;
  h00 = 2 * s3 - 3 * s2 + 1;
  h10 = s3 - 2 * s2 + s;
  h01 = -2 * s3 + 3 * s2;
  h11 = s3 - s2;
}

__device__ __forceinline__ double hermite_z(double p0, double v0, double p1, double v1, double dt, double s) {
  double h00, h10, h01, h11;
  hbasis(s, h00, h10, h01, h11);
  return h00 * p0 + h10 * dt * v0 + h01 * p1 + h11 * dt * v1;
}

struct ScreenParams {
  int n;
  double z[SPT_MAX_SCREENS];
};

// Screen crossings of one particle over one accepted step: cubic Hermite
// z(s), split at its real extrema, half-open (t_old, t_new] sign test on
// moving particles, 60-step bisection (reference screens.py:43-168).
__device__ __forceinline__ void screens_particle_body(const ScreenParams& SP, spt_event* events, int64_t cap,
                                                     unsigned long long* count, int group, int particle, int64_t batch,
                                                     double t0, double t1, const double* p0, const double* v0,
                                                     const double* p1, const double* v1) {
  if (SP.n == 0) return;
  const double dt = t1 - t0;
  const bool moving = p0[0] != p1[0] || p0[1] != p1[1] || p0[2] != p1[2] || v0[0] != v1[0] || v0[1] != v1[1] ||
                      v0[2] != v1[2];
  if (!moving || !(dt > 0.0)) return;
  const double z0 = p0[2], z1 = p1[2];
  const double u0 = dt * v0[2], u1 = dt * v1[2];
  double A = 3 * (2 * (z0 - z1) + u0 + u1);
  double B = 2 * (3 * (z1 - z0) - 2 * u0 - u1);
  double C = u0;
  double sc = fmax(fmax(fabs(A), fabs(B)), fabs(C));
  sc = sc > 0.0 ? sc : 1.0;
  A = A / sc;
  B = B / sc;
  C = C / sc;
  const double disc = B * B - 4 * A * C;
  const double q = -.5 * (B + (B >= 0.0 ? 1.0 : -1.0) * sqrt(fmax(disc, 0.0)));
  double r0 = q / (A != 0.0 ? A : 1.0);
  const double r1 = C / (q != 0.0 ? q : 1.0);
  const double lin = -C / (B != 0.0 ? B : 1.0);
  if (A == 0.0) r0 = lin;
  const bool ok0 = (disc >= 0.0) && (A != 0.0 || B != 0.0) && (r0 > 0.0) && (r0 < 1.0);
  const bool ok1 = (disc >= 0.0) && (A != 0.0) && (q != 0.0) && (r1 > 0.0) && (r1 < 1.0);
  const double ca = ok0 ? r0 : 1.0, cb = ok1 ? r1 : 1.0;
  const double cuts[4] = {0.0, fmin(ca, cb), fmax(ca, cb), 1.0};
  for (int si = 0; si < SP.n; ++si) {
    const double zs = SP.z[si];
    for (int iv = 0; iv < 3; ++iv) {
      const double left = cuts[iv], right = cuts[iv + 1];
      if (!(right > left)) continue;
      const double f0 = hermite_z(z0, v0[2], z1, v1[2], dt, left) - zs;
      const double f1 = hermite_z(z0, v0[2], z1, v1[2], dt, right) - zs;
      if (!(f0 != 0.0) || !(sgn(f0) != sgn(f1))) continue;
      double lo = left, hi = right, flo = f0;
      for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double fm = hermite_z(z0, v0[2], z1, v1[2], dt, mid) - zs;
        if (sgn(fm) == sgn(flo)) {
          lo = mid;
          flo = fm;
        } else {
          hi = mid;
        }
      }
      const double ss = 0.5 * (lo + hi);
      double h00, h10, h01, h11;
      hbasis(ss, h00, h10, h01, h11);
      const double s2 = ss * ss;
      const double d00 = 6 * s2 - 6 * ss, d10 = 3 * s2 - 4 * ss + 1, d01 = -6 * s2 + 6 * ss, d11 = 3 * s2 - 2 * ss;
      const unsigned long long k = atomicAdd(count, 1ULL);
      if ((int64_t)k < cap) {
        spt_event* e = events + k;
        e->group = group;
        e->particle = particle;
        e->screen = si;
        e->interval = iv;
        e->batch = batch;
        e->t = t0 + ss * dt;
        for (int c = 0; c < 3; ++c) {
          e->pos[c] = h00 * p0[c] + (h10 * dt) * v0[c] + h01 * p1[c] + (h11 * dt) * v1[c];
          e->vel[c] = (d00 / dt) * p0[c] + d10 * v0[c] + (d01 / dt) * p1[c] + d11 * v1[c];
        }
      }
    }
  }
}

// Out of line for the persistent kernel (its register budget); the
// lockstep end kernel inlines the body instead (an out-of-line call there
// spilled the caller's registers to 1.6 KB of scratch per lane).
__device__ __noinline__ void screens_particle(const ScreenParams& SP, spt_event* events, int64_t cap,
                                              unsigned long long* count, int group, int particle, int64_t batch,
                                              double t0, double t1, const double* p0, const double* v0,
                                              const double* p1, const double* v1) {
  screens_particle_body(SP, events, cap, count, group, particle, batch, t0, t1, p0, v0, p1, v1);
}

// ------------------------------------------------------------ trace kernel
#include "spt_toroidal.inc"
#include "spt_image.inc"

struct TraceParams {
  spt_geometry G;
  BemView bem;
  ScreenParams SP;
  int n_groups;
  int64_t n_total;
  const int64_t* offsets;
  const int32_t* order;
  double* pos;
  double* vel;
  const double* t_birth;
  const double* charge;
  const double* mass;
  double* t_death;
  double* kbuf;
  double q_over_m, a2, t_max, rtol, atol, safety, fmin_, fmax_, init_frac;
  int64_t max_steps;
  int64_t* attempts;
  int64_t* accepted;
  int64_t* active_stages;
  double* margin;
  int32_t* status;
  spt_event* events;
  int64_t ev_cap;
  unsigned long long* ev_count;
  int* queue;
  const double* t_out;
  int n_out;
  spt_event* snaps;
  int64_t snap_cap;
  unsigned long long* snap_count;
  double* t_end;  // per group: t after the last attempt, or nullptr
  ImageView IB;   // image-charge BEM of the persistent kernel's image instantiations
};

// Output times at or before t (the ascending list's count below t's
// position), and one trajectory record (reference batched_rk45.py:237-242,
// trajectories.py:44-74).
__device__ __forceinline__ int outputs_reached(const TraceParams& P, double t) {
  int k = 0;
  while (k < P.n_out && P.t_out[k] <= t) ++k;
  return k;
}

__device__ __forceinline__ void record_snapshot(const TraceParams& P, int k, int g, int p, int64_t batch, double t,
                                                const double* x, const double* v) {
  const unsigned long long i = atomicAdd(P.snap_count, 1ull);
  if (i >= (unsigned long long)P.snap_cap) return;
  spt_event& e = P.snaps[i];
  e.group = g;
  e.particle = p;
  e.screen = k;
  e.interval = 0;
  e.batch = batch;
  e.t = t;
  for (int c = 0; c < 3; ++c) {
    e.pos[c] = x[c];
    e.vel[c] = v[c];
  }
}

// A group's start time coinciding exactly with an output time records its
// particles born by then (reference batched_rk45.py:189-209).
__device__ __forceinline__ void snapshot_at_start(const TraceParams& P, int g, int p, int64_t gi, double t, double tb) {
  if (P.n_out == 0) return;
  const int last = outputs_reached(P, t) - 1;
  if (last >= 0 && P.t_out[last] == t && tb <= t) {
    const double x[3] = {P.pos[3 * gi], P.pos[3 * gi + 1], P.pos[3 * gi + 2]};
    const double v[3] = {P.vel[3 * gi], P.vel[3 * gi + 1], P.vel[3 * gi + 2]};
    record_snapshot(P, last, g, p, -1, t, x, v);
  }
}

template <int T, int NMAX>
struct TraceShared {
  double sx[NMAX], sy[NMAX], sz[NMAX], sq[NMAX];
  double fx[NMAX], fy[NMAX], fz[NMAX];
  int evalid[NMAX];  // scratch slots 48..50 hold the external field at the state
  int scomp[NMAX];  // field to be evaluated at this stage
  double cx[NMAX], cy[NMAX], cz[NMAX];
  double ox[NMAX], oy[NMAX], oz[NMAX];
  int sact[NMAX];
  int cidx[NMAX];
  int need[NMAX];
  double red[T];
  int slot;
  int m;
  BemShared<T, NMAX> bem;
};

// External field of every active particle of the group at the stage
// positions in LDS (sx, sy, sz) into (fx, fy, fz). Stage 1 sits at the
// state, whose field is known after the first attempt (the previous
// attempt's stage-7 field after an accepted step, its stage-1 field after
// a rejected one): copied instead of evaluated. The field depends on the
// position only; the work count stays 7 stages per attempt.
template <int T, int NMAX, int PR>
__device__ void stage_field(const TraceParams& P, int64_t off, int n, int st, TraceShared<T, NMAX>& sh) {
  const int tid = threadIdx.x;
  if (tid < n) {
    const bool known = sh.sact[tid] && st == 0 && sh.evalid[tid];
    sh.scomp[tid] = sh.sact[tid] && !known;
    if (known) {
      const size_t nt = (size_t)P.n_total, gi = (size_t)(off + tid);
      sh.fx[tid] = P.kbuf[48 * nt + gi];
      sh.fy[tid] = P.kbuf[49 * nt + gi];
      sh.fz[tid] = P.kbuf[50 * nt + gi];
    }
  }
  if (!is_bem(P.G)) {
    if (tid < n && sh.scomp[tid])
      geometry_field_point(P.G, sh.sx[tid], sh.sy[tid], sh.sz[tid], 0, 0, 0, sh.fx[tid], sh.fy[tid], sh.fz[tid]);
    __syncthreads();
    return;
  }
  if (tid < n) {
    bool need = sh.scomp[tid] && bem_needed(P.G, sh.sx[tid], sh.sy[tid], sh.sz[tid]);
    double qx, qy, qz;
    if (need && P.bem.mp && multipole_pert(P.bem, sh.sx[tid], sh.sy[tid], sh.sz[tid], qx, qy, qz)) {
      geometry_field_point(P.G, sh.sx[tid], sh.sy[tid], sh.sz[tid], qx, qy, qz, sh.fx[tid], sh.fy[tid], sh.fz[tid]);
      sh.scomp[tid] = 0;
      need = false;
    }
    sh.need[tid] = need;
  }
  __syncthreads();
  if (tid == 0) {
    int m = 0;
    for (int i = 0; i < n; ++i)
      if (sh.need[i]) sh.cidx[m++] = i;
    sh.m = m;
  }
  __syncthreads();
  const int m = sh.m;
  if (tid < m) {
    const int i = sh.cidx[tid];
    sh.cx[tid] = sh.sx[i];
    sh.cy[tid] = sh.sy[i];
    sh.cz[tid] = sh.sz[i];
  }
  __syncthreads();
  bem_pert_any<T, NMAX, PR>(P.bem, m, sh.cx, sh.cy, sh.cz, sh.ox, sh.oy, sh.oz, sh.bem);
  if (tid < m) {
    const int i = sh.cidx[tid];
    geometry_field_point(P.G, sh.sx[i], sh.sy[i], sh.sz[i], sh.ox[tid], sh.oy[tid], sh.oz[tid], sh.fx[i], sh.fy[i],
                         sh.fz[i]);
  }
  if (tid < n && sh.scomp[tid] && !sh.need[tid])
    geometry_field_point(P.G, sh.sx[tid], sh.sy[tid], sh.sz[tid], 0, 0, 0, sh.fx[tid], sh.fy[tid], sh.fz[tid]);
  __syncthreads();
}

// Per-particle scratch slots (coalesced across a group's threads):
// 0..41  k of stage st, component c (0-2 position, 3-5 velocity): st*6+c
// 42..47 candidate end-of-step position and velocity
// 48..50 external field at the state; 51..53 at the candidate end point
#define KB(sl) P.kbuf[(size_t)(sl) * (size_t)P.n_total + (size_t)gi]
constexpr int kSlots = 54;

// Integrate one group from its earliest birth to t_max (reference
// batched_rk45.py:149-323 for a single group; _attempt_step 326-401).
// Particle state lives in global memory (caller's pos/vel plus the stage
// scratch) so that only a handful of scalars are live across the field
// gather.
// Per-workgroup LDS of the image-charge BEM (NM > 0 modes in registers;
// NM = 0: none).
template <int NM>
union ImgLds {
  ImgRhsLds<NM> r;
  ImgForceLds f;
};
template <int NM>
struct ImgWork {
  ImgLds<NM> u;
};
template <>
struct ImgWork<0> {};

// PR: SPT_PREC_FP64, SPT_PREC_MIXED (FP32 far-field sums and pairs) or
// SPT_PREC_PAIR32 (FP64 field, FP32 + Kahan pairs). NM > 0: the
// image-charge BEM force of the group at every stage (T = NMAX = 256).
template <int T, int NMAX, int PR, int NM = 0>
__device__ void run_group(const TraceParams& P, int g, TraceShared<T, NMAX>& sh, ImgWork<NM>& iw) {
  constexpr int FPR = PR == SPT_PREC_MIXED ? SPT_PREC_MIXED : SPT_PREC_FP64;  // field
  constexpr int APR = PR == SPT_PREC_FP64 ? SPT_PREC_FP64 : SPT_PREC_MIXED;   // pairs
  static_assert(NM == 0 || (T == 256 && NMAX == 256), "the image-charge BEM runs in 256-thread workgroups");
  const int tid = threadIdx.x;
  const int64_t off = P.offsets[g];
  const int n = (int)(P.offsets[g + 1] - off);
  const bool mine = tid < n;
  const int64_t gi = off + tid;
  double tb = INFINITY, td = INFINITY;
  if (mine) {
    tb = P.t_birth[gi];
    sh.sq[tid] = P.charge[gi];
    sh.evalid[tid] = 0;
  }
  double t = block_min<T>(mine ? tb : INFINITY, sh.red);
  double h = (P.t_max - t) * P.init_frac;
  if (mine) snapshot_at_start(P, g, tid, gi, t, tb);
  bool running = true;
  int64_t attempts = 0, nacc = 0, act_st = 0;
  double margin = INFINITY;
  int status = SPT_STATUS_DONE;
  const double tol = 1e-9 * fmax(P.t_max, 1e-30);
  const double ncomp = 6.0 * n;
  while (running) {
    if (attempts >= P.max_steps) {
      status = SPT_STATUS_STEPCAP;
      break;
    }
    // next birth checkpoint: smallest t_birth > t (reference 228-233)
    const double n_reached = block_sum<T>((mine && tb <= t) ? 1.0 : 0.0, sh.red);
    const double nxt_cp = block_min<T>((mine && tb > t) ? tb : INFINITY, sh.red);
    const bool exhausted = n_reached >= (double)n;
    const double next = exhausted ? P.t_max : nxt_cp;
    const int optr = outputs_reached(P, t);
    const bool oexh = optr >= P.n_out;
    const double nout = oexh ? P.t_max : P.t_out[optr];
    const double target = fmin(fmin(next, P.t_max), nout);
    const double hr = fmin(h, target - t);
    const bool active = mine && (t >= tb) && (t < td);
    const double n_act = block_sum<T>(active ? 1.0 : 0.0, sh.red);
#pragma unroll 1
    for (int st = 0; st < 7; ++st) {
      if (mine) {
        double ps[3], vs[3];
        if (st < 6) {
          for (int c = 0; c < 3; ++c) {
            ps[c] = P.pos[3 * gi + c];
            vs[c] = P.vel[3 * gi + c];
          }
          for (int j = 0; j < st; ++j) {
            const double ha = hr * c_A[st][j];
            for (int c = 0; c < 3; ++c) {
              ps[c] = ps[c] + ha * KB(j * 6 + c);
              vs[c] = vs[c] + ha * KB(j * 6 + 3 + c);
            }
          }
        } else {
          for (int c = 0; c < 3; ++c) {
            ps[c] = KB(42 + c);
            vs[c] = KB(45 + c);
          }
        }
        sh.sx[tid] = ps[0];
        sh.sy[tid] = ps[1];
        sh.sz[tid] = ps[2];
        sh.sact[tid] = active ? 1 : 0;
        for (int c = 0; c < 3; ++c) KB(st * 6 + c) = active ? vs[c] : 0.0;
      }
      __syncthreads();
      stage_field<T, NMAX, FPR>(P, off, n, st, sh);
      if constexpr (NM > 0) {  // the image-charge force at the stage positions, into cx, cy, cz
        auto mem = [&](int e, double& x, double& y, double& z, double& q) {
          x = sh.sx[e];
          y = sh.sy[e];
          z = sh.sz[e];
          q = sh.sact[e] ? sh.sq[e] : 0.0;
        };
        __syncthreads();
        img_rhs_group<NM>(P.IB, off, n, g, mem, iw.u.r);
        __syncthreads();
        img_solve_group(P.IB, g);
        __syncthreads();
        auto memf = [&](int e, double& x, double& y, double& z, double& q, int& a) {
          x = sh.sx[e];
          y = sh.sy[e];
          z = sh.sz[e];
          a = sh.sact[e];
          q = a ? sh.sq[e] : 0.0;
        };
        auto out = [&](int e, double fx, double fy, double fz) {
          sh.cx[e] = fx;
          sh.cy[e] = fy;
          sh.cz[e] = fz;
        };
        img_force_group<NM>(P.IB, P.G, off, n, g, P.a2, memf, out, iw.u.f);
        __syncthreads();
      }
      if (mine && active) {  // keep stage 1's field (the state's) and stage 7's (the end point's)
        if (st == 0) {
          KB(48) = sh.fx[tid];
          KB(49) = sh.fy[tid];
          KB(50) = sh.fz[tid];
          sh.evalid[tid] = 1;
        } else if (st == 6) {
          KB(51) = sh.fx[tid];
          KB(52) = sh.fy[tid];
          KB(53) = sh.fz[tid];
        }
      }
      if (mine) {
        double Fc[3], Fi[3];
        pair_forces_p<APR>(P.G, tid, n, sh.sx, sh.sy, sh.sz, sh.sq, sh.sact, P.a2, Fc, Fi);
        if constexpr (NM > 0) {
          Fi[0] = sh.cx[tid];
          Fi[1] = sh.cy[tid];
          Fi[2] = sh.cz[tid];
        }
        const double mass = P.mass[gi];
        const double F[3] = {sh.fx[tid], sh.fy[tid], sh.fz[tid]};
        for (int c = 0; c < 3; ++c) {
          const double ac = n == 1 ? 0.0 : Fc[c] / mass;
          const double a = P.q_over_m * F[c] + ac + Fi[c] / mass;
          KB(st * 6 + 3 + c) = active ? a : 0.0;
        }
        if (st == 5) {
          for (int c = 0; c < 3; ++c) {
            double sp = c_B[0] * KB(c), sv = c_B[0] * KB(3 + c);
            for (int k = 1; k < 6; ++k) {
              sp = sp + c_B[k] * KB(k * 6 + c);
              sv = sv + c_B[k] * KB(k * 6 + 3 + c);
            }
            KB(42 + c) = P.pos[3 * gi + c] + hr * sp;
            KB(45 + c) = P.vel[3 * gi + c] + hr * sv;
          }
        }
      }
      __syncthreads();
    }
    double lp = 0.0, lv = 0.0;
    if (mine) {
      for (int c = 0; c < 3; ++c) {
        double ep = c_E[0] * KB(c), ev = c_E[0] * KB(3 + c);
        for (int k = 1; k < 7; ++k) {
          ep = ep + c_E[k] * KB(k * 6 + c);
          ev = ev + c_E[k] * KB(k * 6 + 3 + c);
        }
        ep = hr * ep;
        ev = hr * ev;
        const double sp = P.atol + P.rtol * fmax(fabs(P.pos[3 * gi + c]), fabs(KB(42 + c)));
        const double sv = kLight * P.atol + P.rtol * fmax(fabs(P.vel[3 * gi + c]), fabs(KB(45 + c)));
        const double a = ep / sp, b = ev / sv;
        lp += a * a;
        lv += b * b;
      }
    }
    const double sq = block_sum<T>(lp, sh.red) + block_sum<T>(lv, sh.red);
    double err = sqrt(sq / ncomp);
    if (err == 0.0) err = 2.2250738585072014e-308;
    const bool acc = err <= 1.0;
    double fac = P.safety * pow(err, -1.0 / 5);
    fac = fmin(fmax(fac, P.fmin_), P.fmax_);
    const int64_t batch = attempts;
    attempts += 1;
    margin = fmin(margin, fabs(err - 1.0));
    act_st += 7 * (int64_t)n_act;
    h = hr * fac;
    if (acc) {
      nacc += 1;
      const double t_old = t;
      t = t + hr;
      if (!exhausted && t >= next - tol) h = (P.t_max - t) * P.init_frac;
      if (mine) {
        double x0[3], v0[3], x1[3], v1[3];
        for (int c = 0; c < 3; ++c) {
          x0[c] = P.pos[3 * gi + c];
          v0[c] = P.vel[3 * gi + c];
          x1[c] = KB(42 + c);
          v1[c] = KB(45 + c);
          P.pos[3 * gi + c] = x1[c];
          P.vel[3 * gi + c] = v1[c];
        }
        if (active) {  // stage 7 sat at the new state
          for (int c = 0; c < 3; ++c) KB(48 + c) = KB(51 + c);
          sh.evalid[tid] = 1;
        }
        if (active && kill_point(P.G, x1[0], x1[1], x1[2])) td = t;
        screens_particle(P.SP, P.events, P.ev_cap, P.ev_count, g, tid, batch, t_old, t, x0, v0, x1, v1);
        if (!oexh && t >= nout - tol && tb <= t) record_snapshot(P, optr, g, tid, batch, t, x1, v1);
      }
      running = t < P.t_max * (1 - 1e-14);
    }
    if (step_stalled(acc, h, t)) {
      status = SPT_STATUS_STALLED;
      break;
    }
  }
  if (mine) P.t_death[gi] = td;
  if (tid == 0) {
    if (P.t_end) P.t_end[g] = t;
    P.attempts[g] = attempts;
    P.accepted[g] = nacc;
    P.active_stages[g] = act_st;
    P.margin[g] = margin;
    P.status[g] = status;
  }
  __syncthreads();
}

template <int T, int NMAX, int PR, int NM = 0>
__global__ __launch_bounds__(T) void trace_kernel(TraceParams P) {
  __shared__ TraceShared<T, NMAX> sh;
  __shared__ ImgWork<NM> iw;
  while (true) {
    if (threadIdx.x == 0) sh.slot = atomicAdd(P.queue, 1);
    __syncthreads();
    const int q = sh.slot;
    __syncthreads();
    if (q >= P.n_groups) break;
    const int g = P.order ? P.order[q] : q;
    run_group<T, NMAX, PR, NM>(P, g, sh, iw);
  }
}



__global__ void reset_int_kernel(int* p, int v) { *p = v; }
__global__ void reset_u64_kernel(unsigned long long* p) { *p = 0ULL; }
__global__ void reset_u32_kernel(unsigned int* p) { *p = 0u; }
__global__ void fill_u32_kernel(unsigned* p, int n, unsigned v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ----------------------------------------------- standalone field kernels
template <int T, int NMAX>
struct FieldShared {
  double px[NMAX], py[NMAX], pz[NMAX];
  double ox[NMAX], oy[NMAX], oz[NMAX];
  int cidx[NMAX];
  int need[NMAX];
  int tile[NMAX];  // binned gather: each point's tile from the packing (-1: none)
  int m;
  BemShared<T, NMAX> bem;
};

// evaluate_axisymmetric_field over PB points per block.
template <int T, int NMAX, int PR>
__global__ __launch_bounds__(T, SPT_FIELD_MIN_WAVES) void bem_pert_kernel(BemView B, const double* pos, int64_t n, double* E,
                                                                          int PB) {
  __shared__ FieldShared<T, NMAX> sh;
  const int64_t base = (int64_t)blockIdx.x * PB;
  const int m_all = (int)min((int64_t)PB, n - base);
  const int tid = threadIdx.x;
  // Points in the multipole zone or in a table-only tile (if any) are done
  // here, as the binned lockstep gather does them; the rest go to the
  // gather in their original order.
  if (tid < m_all) {
    const double x = pos[3 * (base + tid)], y = pos[3 * (base + tid) + 1], z = pos[3 * (base + tid) + 2];
    double qx, qy, qz;
    bool far = B.mp && multipole_pert(B, x, y, z, qx, qy, qz);
    if (!far) {
      double cx, cz, h;
      const int t = tile_only(B, x, y, z, cx, cz, h);
      if (t >= 0) {
        tile_point(B, t, x, y, z, cx, cz, h, qx, qy, qz);
        far = true;
      }
    }
    if (far) {
      E[3 * (base + tid)] = qx;
      E[3 * (base + tid) + 1] = qy;
      E[3 * (base + tid) + 2] = qz;
    }
    sh.need[tid] = !far;
  }
  __syncthreads();
  if (tid == 0) {
    int m = 0;
    for (int i = 0; i < m_all; ++i)
      if (sh.need[i]) sh.cidx[m++] = i;
    sh.m = m;
  }
  __syncthreads();
  const int m = sh.m;
  if (tid < m) {
    const int64_t gi = base + sh.cidx[tid];
    sh.px[tid] = pos[3 * gi];
    sh.py[tid] = pos[3 * gi + 1];
    sh.pz[tid] = pos[3 * gi + 2];
  }
  __syncthreads();
  bem_pert_any<T, NMAX, PR>(B, m, sh.px, sh.py, sh.pz, sh.ox, sh.oy, sh.oz, sh.bem);
  if (tid < m) {
    const int64_t gi = base + sh.cidx[tid];
    E[3 * gi] = sh.ox[tid];
    E[3 * gi + 1] = sh.oy[tid];
    E[3 * gi + 2] = sh.oz[tid];
  }
}

// Geometry.field over PB points per block (BEM part compacted to the
// points that need it).
template <int T, int NMAX, int PR>
__global__ __launch_bounds__(T, SPT_FIELD_MIN_WAVES) void geometry_field_kernel(spt_geometry G, BemView B,
                                                                                const double* pos, int64_t n,
                                                                                double* E, int PB) {
  __shared__ FieldShared<T, NMAX> sh;
  __shared__ double sx[NMAX], sy[NMAX], sz[NMAX];
  const int64_t base = (int64_t)blockIdx.x * PB;
  const int m_all = (int)min((int64_t)PB, n - base);
  const int tid = threadIdx.x;
  if (tid < m_all) {
    sx[tid] = pos[3 * (base + tid)];
    sy[tid] = pos[3 * (base + tid) + 1];
    sz[tid] = pos[3 * (base + tid) + 2];
    bool need = is_bem(G) && bem_needed(G, sx[tid], sy[tid], sz[tid]);
    double qx, qy, qz;
    if (need && B.mp && multipole_pert(B, sx[tid], sy[tid], sz[tid], qx, qy, qz)) {
      double ex, ey, ez;
      geometry_field_point(G, sx[tid], sy[tid], sz[tid], qx, qy, qz, ex, ey, ez);
      E[3 * (base + tid)] = ex;
      E[3 * (base + tid) + 1] = ey;
      E[3 * (base + tid) + 2] = ez;
      need = false;
      sx[tid] = NAN;  // done: skipped below
    }
    sh.need[tid] = need;
  }
  __syncthreads();
  if (tid == 0) {
    int m = 0;
    for (int i = 0; i < m_all; ++i)
      if (sh.need[i]) sh.cidx[m++] = i;
    sh.m = m;
  }
  __syncthreads();
  const int m = sh.m;
  if (tid < m) {
    const int i = sh.cidx[tid];
    sh.px[tid] = sx[i];
    sh.py[tid] = sy[i];
    sh.pz[tid] = sz[i];
  }
  __syncthreads();
  bem_pert_any<T, NMAX, PR>(B, m, sh.px, sh.py, sh.pz, sh.ox, sh.oy, sh.oz, sh.bem);
  if (tid < m) {
    const int i = sh.cidx[tid];
    double ex, ey, ez;
    geometry_field_point(G, sx[i], sy[i], sz[i], sh.ox[tid], sh.oy[tid], sh.oz[tid], ex, ey, ez);
    E[3 * (base + i)] = ex;
    E[3 * (base + i) + 1] = ey;
    E[3 * (base + i) + 2] = ez;
  }
  if (tid < m_all && !sh.need[tid] && !isnan(sx[tid])) {
    double ex, ey, ez;
    geometry_field_point(G, sx[tid], sy[tid], sz[tid], 0, 0, 0, ex, ey, ez);
    E[3 * (base + tid)] = ex;
    E[3 * (base + tid) + 1] = ey;
    E[3 * (base + tid) + 2] = ez;
  }
}

#include "spt_mlp.inc"

// ============================================================ lockstep mode
// The reference's schedule (batched_rk45.py:219-319): every running group
// makes one attempt per iteration. Per attempt: 7 x (field gather over all
// particles, then per-group accel), then one per-group end-of-attempt
// kernel that also prepares the next attempt. Group scalars and particle
// stage data live in global scratch between kernels.
//
// Extra per-particle slots: 48..50 stage position, 51..53 field there,
// 54..56 the external field at the particle's current state, 57..59 the
// field at this attempt's end point (stage 7).
//
// Stage-1 field reuse: the first stage of an attempt sits at the state,
// where the field is already known (after an accepted step the state is
// the previous attempt's stage-7 point; after a rejected one, the previous
// attempt's stage-1 point) and the external field depends on the position
// only, so it is copied instead of gathered again. Pair and image forces,
// which depend on the frozen active set, are recomputed at every stage.
// The work count stays 7 stages per attempt.
constexpr int kSlotsLockstep = 60;
constexpr int LS_RUNNING = 1, LS_EXHAUSTED = 2;

// Binned gather of far-rule lockstep traces. At the start of every attempt
// the particles whose state needs the BEM perturbation are counted into
// bins (bin_key of their last gather: distance band, closest segment; in
// the multipole zone, bins of log r) and given consecutive slots bin by
// bin; at every stage the kernel that computes the stage position writes it
// into the particle's slot record, the gather kernel reads 64 consecutive
// records per block (one bin or neighbouring bins: the lanes of a wavefront
// need the same segments, rules, expansions or multipole terms) and writes
// the perturbation and the new bin into the same slot, and the stage's
// acceleration kernel reads it back. Slot order within a bin comes from
// atomics; a far-rule gather's result does not depend on the point's
// block-mates, so traces stay bit-reproducible.
constexpr int kBinMaxKeys = 8192;
constexpr int kBinThreads = 1024;
struct BinView {
  int32_t* key;      // per particle: bin of its last gather (nullptr: binning off)
  int32_t* slot;     // per particle: its slot this attempt, -1 none (-2: eligible, between the count and slot kernels)
  uint8_t* src;      // per particle, this stage: 1 = the field comes from its slot's output record
  unsigned* hist;    // n_keys: eligible particles per bin (zero between attempts)
  unsigned* cursor;  // n_keys: next free slot per bin
  unsigned* total;   // slots in use this attempt (grows by particles that need a gather without a slot)
  double2* rin;      // per slot: stage position (rho, z)
  int32_t* rtag;     // per slot: record kind (kRecGather, or the other kRec* tags)
  double2* rout;     // per slot: perturbation (E_rho, E_z)
  int32_t* rkey;     // per slot: the point's bin
  int n_keys, nsk, shift;
  int key_tile0;     // first tile bin (tile t: key_tile0 + t), after the distance-band bins
  int key_mp0;       // first multipole-zone bin (n_keys when the zone is off)
  int append_only;   // 1: no slots at the start of an attempt, every record through the append path
  int32_t* glist;    // the slots of this stage's gather records, in no particular order
  unsigned* gcount;  // per stage index 0..6: records in glist (kListOff: the list is off this attempt)
  unsigned* list_on; // this attempt fills glist (set with the counts by the scan kernel)
  unsigned* n_start; // eligible particles in gather bins at the attempt's start (the count kernel's)
  int key_tab0;      // first bin after the gather bins (the table-only tiles' and multipole bins follow)
};
// At most this many gather records: the few-records form (16 points per
// block, 16 threads each; chosen from the count alone, so results do not
// depend on slot order).
constexpr unsigned kFewRecords = 8192;
// An attempt's records go into the compact list only when its gather bins
// hold at most kListFrom slots at its start (the point mix changes little
// within an attempt); otherwise every stage's count is kListOff and the
// field kernel takes the slot-order form.
constexpr unsigned kListFrom = kFewRecords / 2;
constexpr unsigned kListOff = 0x40000000u;
constexpr int kBinMpKeys = 64;  // multipole-zone bins: 16 per octave of r^2 beyond r_min, open-ended

struct LockParams {
  TraceParams T;
  const int32_t* pgroup;   // particle -> group
  uint8_t* active;         // per particle, frozen for the current attempt
  double* g_t;
  double* g_h;
  double* g_hr;
  double* g_next;
  double* g_nact;
  int32_t* g_flags;
  unsigned int* done;      // finished groups
  uint8_t* sur;            // per particle: surrogate flag of this stage
  uint8_t* fvalid;         // per particle: slots 54..56 hold the field at the current state
  MlpView V;               // the surrogate (when sur != nullptr): split BF16
  Mlp64View V64;           // and FP64 forms
  const double* imgF;      // image-charge BEM force per particle (n x 3), or nullptr
  const int32_t* order;    // ls_field_kernel's particle order (a permutation of 0..n-1), or nullptr
  BinView X;               // the binned gather (X.key == nullptr: ls_field_kernel)
};

// Particle taken by position k of ls_field_kernel's sweep.
__device__ __forceinline__ int64_t ls_field_particle(const LockParams& L, int64_t k) {
  return L.order ? (int64_t)L.order[k] : k;
}

__global__ void ls_pgroup_kernel(const int64_t* offsets, int n_groups, int32_t* pgroup) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_groups) return;
  for (int64_t i = offsets[g]; i < offsets[g + 1]; ++i) pgroup[i] = g;
}

#define LKB(sl) L.T.kbuf[(size_t)(sl) * (size_t)L.T.n_total + (size_t)gi]

// Sum and minimum over one group's lanes: the block (GPB = 1) or, for four
// groups per 64-lane block, the group's 16-lane segment (lane i takes lane
// i + s, s = 8 .. 1: with zero lanes beyond the group, bit-identical to
// block_sum<64>'s s = 32 .. 1 tree, whose first two levels add zeros).
template <int TG, int GPB>
__device__ __forceinline__ double group_sum(double v, double* red) {
  if constexpr (GPB == 1) {
    return block_sum<TG>(v, red);
  } else {
    static_assert(TG / GPB == 16 && TG % 64 == 0, "16-lane groups, four per wavefront");
#pragma unroll
    for (int s = 8; s > 0; s >>= 1) v = v + __shfl_down(v, s, 16);
    return __shfl(v, 0, 16);
  }
}
template <int TG, int GPB>
__device__ __forceinline__ double group_min(double v, double* red) {
  if constexpr (GPB == 1) {
    return block_min<TG>(v, red);
  } else {
    static_assert(TG / GPB == 16 && TG % 64 == 0, "16-lane groups, four per wavefront");
#pragma unroll
    for (int s = 8; s > 0; s >>= 1) v = fmin(v, __shfl_down(v, s, 16));
    return __shfl(v, 0, 16);
  }
}

// Start an attempt for group g at (t, h): birth checkpoint, clipped step,
// frozen active set, stage-0 state (reference batched_rk45.py:223-249,348).
// tid: the thread's member index within the group.
template <int TG, int GPB = 1>
__device__ void ls_begin(const LockParams& L, int g, int n, int tid, double t, double h, int64_t attempts,
                         double* red) {
  const bool mine = tid < n;
  const int64_t gi = L.T.offsets[g] + tid;
  if (attempts >= L.T.max_steps) {
    if (tid == 0) {
      L.T.status[g] = SPT_STATUS_STEPCAP;
      L.g_flags[g] = 0;
      atomicAdd(L.done, 1u);
    }
    return;
  }
  const double tb = mine ? L.T.t_birth[gi] : INFINITY;
  const double n_reached = group_sum<TG, GPB>((mine && tb <= t) ? 1.0 : 0.0, red);
  const double nxt_cp = group_min<TG, GPB>((mine && tb > t) ? tb : INFINITY, red);
  const bool exhausted = n_reached >= (double)n;
  const double next = exhausted ? L.T.t_max : nxt_cp;
  const int optr = outputs_reached(L.T, t);
  const double nout = optr >= L.T.n_out ? L.T.t_max : L.T.t_out[optr];
  const double target = fmin(fmin(next, L.T.t_max), nout);
  const double hr = fmin(h, target - t);
  const bool active = mine && (t >= tb) && (t < L.T.t_death[gi]);
  const double n_act = group_sum<TG, GPB>(active ? 1.0 : 0.0, red);
  if (mine) {
    L.active[gi] = active ? 1 : 0;
    for (int c = 0; c < 3; ++c) LKB(48 + c) = L.T.pos[3 * gi + c];
  }
  if (tid == 0) {
    L.g_hr[g] = hr;
    L.g_next[g] = next;
    L.g_nact[g] = n_act;
    L.g_flags[g] = LS_RUNNING | (exhausted ? LS_EXHAUSTED : 0);
  }
}

template <int TG>
__global__ __launch_bounds__(TG) void ls_init_kernel(LockParams L) {
  __shared__ double red[TG];
  const int g = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t off = L.T.offsets[g];
  const int n = (int)(L.T.offsets[g + 1] - off);
  const bool mine = tid < n;
  const int64_t gi = off + tid;
  if (mine) {
    L.T.t_death[gi] = INFINITY;
    L.fvalid[gi] = 0;
  }
  const double t = block_min<TG>(mine ? L.T.t_birth[gi] : INFINITY, red);
  const double h = (L.T.t_max - t) * L.T.init_frac;
  if (mine) snapshot_at_start(L.T, g, tid, gi, t, L.T.t_birth[gi]);
  if (tid == 0) {
    L.g_t[g] = t;
    L.g_h[g] = h;
    if (L.T.t_end) L.T.t_end[g] = t;
    L.T.attempts[g] = 0;
    L.T.accepted[g] = 0;
    L.T.active_stages[g] = 0;
    L.T.margin[g] = INFINITY;
    L.T.status[g] = SPT_STATUS_DONE;
  }
  __syncthreads();
  ls_begin<TG>(L, g, n, tid, t, h, 0, red);
}

// Keep a freshly evaluated stage field (slots 51..53): stage 1's as the
// field at the state, stage 7's as the field at the attempt's end point.
__device__ __forceinline__ void ls_keep_field(const LockParams& L, int64_t gi, int st) {
  if (st == 0) {
    for (int c = 0; c < 3; ++c) LKB(54 + c) = LKB(51 + c);
    L.fvalid[gi] = 1;
  } else if (st == 6) {
    for (int c = 0; c < 3; ++c) LKB(57 + c) = LKB(51 + c);
  }
}

// ---- binned gather kernels

// Initial bins: the closest segment over all segments at the stage-0
// position (first index wins) and the distance band.
__global__ __launch_bounds__(256) void ls_bin_key_kernel(LockParams L) {
  const int64_t gi = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (gi >= L.T.n_total) return;
  const BemView& B = L.T.bem;
  const double x = LKB(48), y = LKB(49), z = LKB(50);
  const double rho = hypot(x, y);
  if (B.tl_tree) {
    double cx, cz, h;
    const int t = tile_find(B, rho, z, cx, cz, h);
    if (t >= 0) {
      L.X.key[gi] = L.X.key_tile0 + t;
      return;
    }
  }
  double bd2 = INFINITY;
  int bseg = 0;
  for (int j = 0; j < B.n_seg; ++j) {
    const double d2 = seg_dist2(B.segs + (size_t)j * SEG_STRIDE, rho, z);
    if (d2 < bd2) {
      bd2 = d2;
      bseg = j;
    }
  }
  L.X.key[gi] = bin_key(bseg, bd2, B.segs[(size_t)bseg * SEG_STRIDE + S_L], L.X.nsk, L.X.shift);
}

// Does the particle's state (the stage-0 position) take a slot this attempt?
__device__ __forceinline__ bool bin_eligible(const LockParams& L, int64_t gi, double x, double y, double z) {
  if (!(L.g_flags[L.pgroup[gi]] & LS_RUNNING) || !L.active[gi]) return false;
  return bem_needed(L.T.G, x, y, z);
}

// Bin of a point in the multipole zone (r^2 >= mp_rmin2).
__device__ __forceinline__ int bin_key_mp(const BinView& X, const BemView& B, double r2) {
  const double u = 16.0 * log2(r2 / B.mp_rmin2);
  const int k = u >= (double)(kBinMpKeys - 1) ? kBinMpKeys - 1 : (u > 0.0 ? (int)u : 0);
  return X.key_mp0 + k;
}

// Eligible particles per bin (block histogram in LDS, then one global atomic
// per bin and block).
__global__ __launch_bounds__(kBinThreads) void ls_bin_count_kernel(LockParams L) {
  __shared__ unsigned cnt[kBinMaxKeys];
  __shared__ unsigned n_gather;  // eligible particles whose bin is a gather bin
  const BinView& X = L.X;
  for (int k = threadIdx.x; k < X.n_keys; k += kBinThreads) cnt[k] = 0u;
  if (threadIdx.x == 0) n_gather = 0u;
  __syncthreads();
  const int64_t gi = (int64_t)blockIdx.x * kBinThreads + threadIdx.x;
  if (gi < L.T.n_total) {
    const bool el = bin_eligible(L, gi, LKB(48), LKB(49), LKB(50));
    const bool e = !X.append_only && el;
    X.slot[gi] = e ? -2 : -1;
    if (e) atomicAdd(&cnt[X.key[gi]], 1u);
    if (el && X.key[gi] < X.key_tab0) atomicAdd(&n_gather, 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < X.n_keys; k += kBinThreads)
    if (cnt[k]) atomicAdd(&X.hist[k], cnt[k]);
  if (threadIdx.x == 0 && n_gather) atomicAdd(X.n_start, n_gather);
}

// First slot of every bin (exclusive prefix sum of the counts, one block),
// the slot total; clears the counts for the next attempt.
__global__ __launch_bounds__(kBinThreads) void ls_bin_scan_kernel(LockParams L) {
  __shared__ unsigned part[kBinThreads];
  const BinView& X = L.X;
  const int tid = threadIdx.x;
  const int per = (X.n_keys + kBinThreads - 1) / kBinThreads;
  const int k0 = tid * per, k1 = min(k0 + per, X.n_keys);
  unsigned own = 0u;
  for (int k = k0; k < k1; ++k) own += X.hist[k];
  part[tid] = own;
  __syncthreads();
  for (int o = 1; o < kBinThreads; o <<= 1) {
    const unsigned v = tid >= o ? part[tid - o] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  unsigned run = part[tid] - own;
  for (int k = k0; k < k1; ++k) {
    const unsigned h = X.hist[k];
    X.cursor[k] = run;
    run += h;
    X.hist[k] = 0u;
  }
  if (tid == kBinThreads - 1) *X.total = part[tid];
  if (tid == 0) {
    // whether this attempt's stages fill the compact list: from the count of
    // particles in gather bins (the same with slots or the append path)
    const bool on = *X.n_start <= kListFrom;
    *X.n_start = 0u;
    *X.list_on = on ? 1u : 0u;
    for (int k = 0; k < 7; ++k) X.gcount[k] = on ? 0u : kListOff;
  }
}

// Slot record tags. Records hold (rho, z): the binned gather works in the
// meridian plane and returns (E_rho, E_z); the acceleration kernel turns
// them around the axis with the stage position's own x / rho and y / rho.
constexpr int kRecGather = 0;     // kRecGather + 1 + t: the gather (64 or 16 records per block), point in
                                  // windowed tile t (t = -1: in no tile)
constexpr int kRecNone = -1;      // no field needed this stage
constexpr int kRecMultipole = -2; // multipole zone, one record per thread
constexpr int kRecTable = -3;     // kRecTable - t: table-only tile t, one record per thread

// The tile a bin key names (key_tile0 + t), or -1: a guess for tile_find_hint.
__device__ __forceinline__ int bin_tile(const BinView& X, int key) {
  const int t = key - X.key_tile0;
  return t >= 0 && key < X.key_mp0 ? t : -1;
}

// Stage st's field for a particle of a running group whose stage position
// (x, y, z) is in slots 48..50: written here where the BEM perturbation is
// not needed (inactive, the state's known field at stage 0, outside the BEM
// region), else the position goes into the particle's slot record (a slot
// at the end of the list for a particle without one), marked for the
// gather, for the multipole zone or for a table-only tile. Every slot of
// the attempt gets its record at every stage. sl: the particle's slot (-1:
// none yet); active: its flag; hint: the tile of its last gather (bin_tile)
// or -1.
__device__ void ls_bin_pack(const LockParams& L, int64_t gi, int st, double x, double y, double z, int sl, bool active,
                            int hint) {
  const BinView& X = L.X;
  int kind = 0;  // 1 gather, 2 multipole zone, 3 table-only tile
  int tile = -1;
  if (active) {
    if (st == 0 && L.fvalid[gi]) {  // the field at the state is known
      for (int c = 0; c < 3; ++c) LKB(51 + c) = LKB(54 + c);
    } else if (bem_needed(L.T.G, x, y, z)) {
      // the zone test on (rho, 0, z), as the field kernel evaluates it
      if (L.T.bem.mp && multipole_zone(L.T.bem, hypot(x, y), 0.0, z)) {
        kind = 2;
      } else {
        double cx, cz, h;
        tile = L.T.bem.tl_tree ? tile_find_hint(L.T.bem, hint, hypot(x, y), z, cx, cz, h) : -1;
        kind = tile >= L.T.bem.tl_nwin ? 3 : 1;  // tiles past tl_nwin: empty windows
      }
    } else {
      double ex, ey, ez;
      geometry_field_point(L.T.G, x, y, z, 0.0, 0.0, 0.0, ex, ey, ez);
      LKB(51) = ex;
      LKB(52) = ey;
      LKB(53) = ez;
      ls_keep_field(L, gi, st);
    }
  }
  // the gather records' slots into the compact list (one atomic per
  // wavefront) when this attempt fills it (a word every thread reads)
  if (*X.list_on) {
    const unsigned long long mask = __ballot(kind == 1);
    if (mask) {
      const int lane = __lane_id();
      const int leader = __ffsll((long long)mask) - 1;
      unsigned base = 0u;
      if (lane == leader) base = atomicAdd(&X.gcount[st], (unsigned)__popcll(mask));
      base = __shfl(base, leader);
      if (kind == 1) {
        if (sl < 0) {
          sl = (int)atomicAdd(X.total, 1u);
          X.slot[gi] = sl;
        }
        X.glist[base + __popcll(mask & ((1ull << lane) - 1ull))] = sl;
      }
    }
  }
  if (kind) {
    if (sl < 0) {
      sl = (int)atomicAdd(X.total, 1u);
      X.slot[gi] = sl;
    }
    X.rin[sl] = make_double2(hypot(x, y), z);
    X.rtag[sl] = kind == 1 ? kRecGather + 1 + tile : (kind == 2 ? kRecMultipole : kRecTable - tile);
  } else if (sl >= 0) {
    X.rtag[sl] = kRecNone;
  }
  X.src[gi] = kind ? 1 : 0;
}

// Slots for the eligible particles, bin by bin (LDS rank within the block,
// one global atomic per bin and block), then stage 0's field or record.
__global__ __launch_bounds__(kBinThreads) void ls_bin_slot_kernel(LockParams L) {
  __shared__ unsigned cnt[kBinMaxKeys];
  const BinView& X = L.X;
  for (int k = threadIdx.x; k < X.n_keys; k += kBinThreads) cnt[k] = 0u;
  __syncthreads();
  const int64_t gi = (int64_t)blockIdx.x * kBinThreads + threadIdx.x;
  const bool mine = gi < L.T.n_total;
  const bool e = mine && X.slot[gi] == -2;
  const int k = e ? X.key[gi] : 0;
  const unsigned r = e ? atomicAdd(&cnt[k], 1u) : 0u;
  __syncthreads();
  for (int kk = threadIdx.x; kk < X.n_keys; kk += kBinThreads)
    if (cnt[kk]) cnt[kk] = atomicAdd(&X.cursor[kk], cnt[kk]);
  __syncthreads();
  if (!mine) return;
  const int sl = e ? (int)(cnt[k] + r) : X.slot[gi];
  if (e) X.slot[gi] = sl;
  if (L.g_flags[L.pgroup[gi]] & LS_RUNNING)
    ls_bin_pack(L, gi, 0, LKB(48), LKB(49), LKB(50), sl, L.active[gi], bin_tile(X, e ? k : X.key[gi]));
}

// Stage st's slot records. Blocks from nb_gather on: the multipole-zone and
// table-only records, one per thread (they sit together after the windowed
// tiles' bins and at the end of the slot list). Blocks below nb_gather: the
// gather records (tag >= 0; the packing's zone and tile tests ran on the
// same position, so the other tags never need the gather), either 64
// consecutive slots per block, lane = slot, or -- with at most kFewRecords
// of them this stage -- 16 entries of the compact list per block, 16
// threads per point: the tail of a trace has few points left near the
// surface, and a lone block's latency then sets the launch's time. The
// choice depends on the count alone (not on slot order), so binned traces
// stay reproducible.
template <int T>
__global__ __launch_bounds__(T, SPT_FIELD_MIN_WAVES) void ls_bin_field_kernel(LockParams L, int nb_gather, int st) {
  __shared__ union {
    FieldShared<T, 64> wide;
    FieldShared<T, 16> few;
  } sh;
  __shared__ int okey[64];
  const BinView& X = L.X;
  const int64_t tot = (int64_t)*X.total;
  const unsigned ng = X.gcount[st];
  const int tid = threadIdx.x;
  if ((int)blockIdx.x >= nb_gather) {
    const int64_t i = (int64_t)(blockIdx.x - nb_gather) * T + tid;
    if (i >= tot) return;
    const int w = X.rtag[i];
    if (w >= kRecNone) return;
    const double2 r = X.rin[i];  // (rho, z): the point at azimuth 0, where E_x = E_rho and E_y = 0
    double qx, qy, qz;
    if (w == kRecMultipole && multipole_pert(L.T.bem, r.x, 0.0, r.y, qx, qy, qz)) {
      X.rout[i] = make_double2(qx, qz);
      X.rkey[i] = bin_key_mp(X, L.T.bem, fma(r.x, r.x, r.y * r.y));
    } else if (w <= kRecTable && kRecTable - w < L.T.bem.tl_n) {
      const int t = kRecTable - w;  // the packing placed the point in this table-only tile
      const double4 c = L.T.bem.tl_cell[t];
      const double h = 0.5 * c.z;
      tile_point(L.T.bem, t, r.x, 0.0, r.y, c.x + h, c.y + h, h, qx, qy, qz);
      X.rout[i] = make_double2(qx, qz);
      X.rkey[i] = X.key_tile0 + t;
    }
    return;
  }
  if (ng <= kFewRecords && ng <= 16u * (unsigned)nb_gather) {
    FieldShared<T, 16>& f = sh.few;
    const unsigned base = blockIdx.x * 16u;
    if (base >= ng) return;
    if (tid < 16) {
      int sl = -1;
      if (base + tid < ng) {
        sl = X.glist[base + tid];
        const double2 r = X.rin[sl];
        f.px[tid] = r.x;
        f.py[tid] = 0.0;
        f.pz[tid] = r.y;
        f.tile[tid] = X.rtag[sl] - kRecGather - 1;
      }
      f.cidx[tid] = sl;
      f.need[tid] = sl >= 0;
    }
    __syncthreads();
    bem_pert_block<T, 16, SPT_PREC_FP64, true, true>(L.T.bem, 16, f.px, f.py, f.pz, f.ox, f.oy, f.oz, f.bem, f.need,
                                                     okey, X.nsk, X.shift, X.key_tile0, f.tile);
    if (tid < 16 && f.need[tid]) {
      X.rout[f.cidx[tid]] = make_double2(f.ox[tid], f.oz[tid]);
      X.rkey[f.cidx[tid]] = okey[tid];
    }
    return;
  }
  FieldShared<T, 64>& f = sh.wide;
  const int64_t base = (int64_t)blockIdx.x * 64;
  if (base >= tot) return;
  const int m = (int)min((int64_t)64, tot - base);
  if (tid < 64) {
    int v = 0;
    const int tg = tid < m ? X.rtag[base + tid] : kRecNone;
    if (tg >= kRecGather) {
      const double2 r = X.rin[base + tid];
      v = 1;
      f.px[tid] = r.x;
      f.py[tid] = 0.0;
      f.pz[tid] = r.y;
      f.tile[tid] = tg - kRecGather - 1;
    }
    f.need[tid] = v;
  }
  if (!__syncthreads_or(tid < 64 && f.need[tid])) return;
  bem_pert_block<T, 64, SPT_PREC_FP64, true, true>(L.T.bem, 64, f.px, f.py, f.pz, f.ox, f.oy, f.oz, f.bem, f.need,
                                                   okey, X.nsk, X.shift, X.key_tile0, f.tile);
  if (tid < m && f.need[tid]) {
    X.rout[base + tid] = make_double2(f.ox[tid], f.oz[tid]);
    X.rkey[base + tid] = okey[tid];
  }
}

// External field at the stage positions of every active particle of every
// running group; PB particles per block, T/PB threads per particle.
template <int T, int NMAX, int PR>
__global__ __launch_bounds__(T, SPT_FIELD_MIN_WAVES) void ls_field_kernel(LockParams L, int PB, int st) {
  __shared__ FieldShared<T, NMAX> sh;
  __shared__ double sx[NMAX], sy[NMAX], sz[NMAX];
  __shared__ int act[NMAX];
  const int64_t base = (int64_t)blockIdx.x * PB;
  const int m_all = (int)min((int64_t)PB, L.T.n_total - base);
  const int tid = threadIdx.x;
  const spt_geometry& G = L.T.G;
  if (tid < m_all) {
    const int64_t gi = ls_field_particle(L, base + tid);
    const int g = L.pgroup[gi];
    act[tid] = (L.g_flags[g] & LS_RUNNING) && L.active[gi];
    sx[tid] = LKB(48);
    sy[tid] = LKB(49);
    sz[tid] = LKB(50);
    if (act[tid] && st == 0 && L.fvalid[gi]) {  // the field at the state is known
      for (int c = 0; c < 3; ++c) LKB(51 + c) = LKB(54 + c);
      act[tid] = 0;
    } else if (act[tid] && L.sur && L.sur[gi] == 1) {  // written by ls_surrogate_kernel
      ls_keep_field(L, gi, st);
      act[tid] = 0;
    }
    bool need = act[tid] && is_bem(G) && bem_needed(G, sx[tid], sy[tid], sz[tid]);
    double qx, qy, qz;
    if (need && L.T.bem.mp && multipole_pert(L.T.bem, sx[tid], sy[tid], sz[tid], qx, qy, qz)) {
      if (L.sur && L.sur[gi] == 2) {  // blend with the parked surrogate perturbation
        const double w = surrogate_weight(L.V.cfg, G, sx[tid], sy[tid], sz[tid]);
        qx = (1.0 - w) * qx + w * LKB(51);
        qy = (1.0 - w) * qy + w * LKB(52);
        qz = (1.0 - w) * qz + w * LKB(53);
      }
      double ex, ey, ez;
      geometry_field_point(G, sx[tid], sy[tid], sz[tid], qx, qy, qz, ex, ey, ez);
      LKB(51) = ex;
      LKB(52) = ey;
      LKB(53) = ez;
      ls_keep_field(L, gi, st);
      need = false;
      act[tid] = 0;
    }
    sh.need[tid] = need;
  }
  __syncthreads();
  if (tid == 0) {
    int m = 0;
    for (int i = 0; i < m_all; ++i)
      if (sh.need[i]) sh.cidx[m++] = i;
    sh.m = m;
  }
  __syncthreads();
  const int m = sh.m;
  if (tid < m) {
    const int i = sh.cidx[tid];
    sh.px[tid] = sx[i];
    sh.py[tid] = sy[i];
    sh.pz[tid] = sz[i];
  }
  __syncthreads();
  bem_pert_any<T, NMAX, PR>(L.T.bem, m, sh.px, sh.py, sh.pz, sh.ox, sh.oy, sh.oz, sh.bem);
  if (tid < m) {
    const int i = sh.cidx[tid];
    const int64_t gi = ls_field_particle(L, base + i);
    double qx = sh.ox[tid], qy = sh.oy[tid], qz = sh.oz[tid];
    if (L.sur && L.sur[gi] == 2) {  // blend with the parked surrogate perturbation
      const double w = surrogate_weight(L.V.cfg, G, sx[i], sy[i], sz[i]);
      qx = (1.0 - w) * qx + w * LKB(51);
      qy = (1.0 - w) * qy + w * LKB(52);
      qz = (1.0 - w) * qz + w * LKB(53);
    }
    double ex, ey, ez;
    geometry_field_point(G, sx[i], sy[i], sz[i], qx, qy, qz, ex, ey, ez);
    LKB(51) = ex;
    LKB(52) = ey;
    LKB(53) = ez;
    ls_keep_field(L, gi, st);
  }
  if (tid < m_all && act[tid] && !sh.need[tid]) {
    const int64_t gi = ls_field_particle(L, base + tid);
    double ex, ey, ez;
    geometry_field_point(G, sx[tid], sy[tid], sz[tid], 0, 0, 0, ex, ey, ez);
    LKB(51) = ex;
    LKB(52) = ey;
    LKB(53) = ez;
    ls_keep_field(L, gi, st);
  }
}

// Surrogate field at the stage positions. Each wavefront takes a run of
// consecutive particles, compacts those of running groups that are active,
// need the BEM perturbation and have a surrogate weight w >= 0, and runs
// the network on them (FP64: 16 per wavefront; three-term BF16 split: 32).
// Flags for ls_field_kernel: 1 = field written (w = 1); 2 = the network's
// perturbation is parked in the field slots and the BEM one is to be
// blended in, (1 - w) BEM + w network; 0 = BEM only.
template <bool F64>
__global__ __launch_bounds__(64) void ls_surrogate_kernel(LockParams L, int st) {
  constexpr int PTS = F64 ? MLP64_PTS : MLP_WAVE_PTS;
  __shared__ union {
    Mlp64WaveShared d;
    MlpWaveShared f;
  } shu;
  double* s_rho = F64 ? shu.d.rho : shu.f.rho;
  double* s_z = F64 ? shu.d.z : shu.f.z;
  double* s_phi = F64 ? shu.d.phi : shu.f.phi;
  double* s_er = F64 ? shu.d.er : shu.f.er;
  double* s_ez = F64 ? shu.d.ez : shu.f.ez;
  int* s_idx = F64 ? shu.d.idx : shu.f.idx;
  const int lane = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * PTS;
  const int cnt = (int)min((int64_t)PTS, L.T.n_total - base);
  const spt_geometry& G = L.T.G;
  bool elig = false;
  double x = 0.0, y = 0.0, z = 0.0, w = -1.0;
  if (lane < cnt) {
    const int64_t gi = base + lane;
    const int g = L.pgroup[gi];
    x = LKB(48);
    y = LKB(49);
    z = LKB(50);
    const bool reuse = st == 0 && L.fvalid[gi];  // ls_field_kernel copies the state's field
    // the multipole zone (exact to its validation tolerance, and cheaper)
    // takes precedence over the network
    const bool mp_zone = L.T.bem.mp && x * x + y * y + z * z >= L.T.bem.mp_rmin2;
    if (!reuse && (L.g_flags[g] & LS_RUNNING) && L.active[gi] && bem_needed(G, x, y, z) && !mp_zone)
      w = surrogate_weight(L.V.cfg, G, x, y, z);
    elig = w >= 0.0;
    L.sur[gi] = !elig ? 0 : (w >= 1.0 ? 1 : 2);
  }
  const unsigned long long mask = __ballot(elig);
  const int m = __popcll(mask);
  if (elig) {
    const int slot = __popcll(mask & ((1ull << lane) - 1ull));
    s_rho[slot] = sqrt(x * x + y * y);
    s_z[slot] = z;
    s_idx[slot] = lane;
  }
  wave_sync();
  if (m == 0) return;
  if (F64)
    mlp_infer64_wave(L.V64, m, s_rho, s_z, shu.d.buf[0], shu.d.buf[1], s_phi, s_er, s_ez);
  else
    mlp_infer_wave(L.V, m, s_rho, s_z, shu.f.buf[0], shu.f.buf[1], s_phi, s_er, s_ez);
  if (lane < m) {
    const int64_t gi = base + s_idx[lane];
    const double px = LKB(48), py = LKB(49), pz = LKB(50);
    const double phi = atan2(py, px);
    const bool axis = s_rho[lane] == 0.0;
    const double er = s_er[lane];
    const double qx = axis ? 0.0 : er * cos(phi), qy = axis ? 0.0 : er * sin(phi), qz = s_ez[lane];
    if (L.sur[gi] == 1) {
      double ex, ey, ez;
      geometry_field_point(G, px, py, pz, qx, qy, qz, ex, ey, ez);
      LKB(51) = ex;
      LKB(52) = ey;
      LKB(53) = ez;
    } else {
      LKB(51) = qx;
      LKB(52) = qy;
      LKB(53) = qz;
    }
  }
}

// Acceleration of stage st for one group (reference geometry.py:289-299,
// batched_rk45.py:362-363), then the next stage's state. GPB groups of at
// most TG / GPB particles per block (GPB > 1 packs small groups so a
// wavefront's lanes are not mostly idle; each group's arithmetic is the
// same as with GPB = 1).
template <int TG, int GPB>
__device__ void ls_end_group(const LockParams& L, int g, int flags, int tid, double* red);

// LAST: stage 7 (st = 6), followed by the end of the attempt in the same
// kernel (ls_end_group) instead of a launch of ls_end_kernel. ST >= 0: the
// stage as a compile-time constant (st is then ignored), so the stage
// combination below keeps only the terms that stage uses; the values are
// the same as with the stage given at run time.
template <int TG, int PR, int GPB = 1, bool LAST = false, int ST = -1>
__global__ __launch_bounds__(TG) void ls_accel_kernel(LockParams L, int st_arg) {
  const int st = ST >= 0 ? ST : st_arg;
  constexpr int W = TG / GPB;  // lanes per group
  __shared__ double sx[TG], sy[TG], sz[TG], sq[TG];
  __shared__ int act[TG];
  const int gs = threadIdx.x / W;
  const int g = blockIdx.x * GPB + gs;
  // the group's flags and extent, loaded together
  const bool in = g < L.T.n_groups;
  const int flags = in ? L.g_flags[g] : 0;
  const int64_t o0 = in ? L.T.offsets[g] : 0, o1 = in ? L.T.offsets[g + 1] : 0;
  const bool run = flags & LS_RUNNING;
  if (GPB == 1 && !run) return;  // block-uniform
  const int tid = threadIdx.x - gs * W;
  const int64_t off = run ? o0 : 0;
  const int n = run ? (int)(o1 - o0) : 0;
  const bool mine = tid < n;
  const int64_t gi = off + tid;
  const int lt = gs * W + tid;  // this thread's LDS slot
  // Everything the stage reads of the particle is loaded up front, so the
  // loads overlap instead of following one another through the kernel.
  const bool binned = L.X.key != nullptr;
  double hr = 0.0, mass = 1.0, p0[3] = {0.0, 0.0, 0.0}, v0[3] = {0.0, 0.0, 0.0}, kprev[5][3];
  bool active = false;
  int sl = -1, hint = -1;
  double fe[3] = {0.0, 0.0, 0.0};  // this stage's field
  if (mine) {
    hr = L.g_hr[g];
    mass = L.T.mass[gi];
    active = L.active[gi];
    if (st < 6) {
      for (int c = 0; c < 3; ++c) {
        p0[c] = L.T.pos[3 * gi + c];
        v0[c] = L.T.vel[3 * gi + c];
      }
#pragma unroll
      for (int j = 0; j < 5; ++j)
        for (int c = 0; c < 3; ++c) kprev[j][c] = j < st ? LKB(j * 6 + 3 + c) : 0.0;
    }
    sx[lt] = LKB(48);
    sy[lt] = LKB(49);
    sz[lt] = LKB(50);
    sq[lt] = L.T.charge[gi];
    act[lt] = active;
    const bool src = binned && L.X.src[gi];
    if (binned) sl = L.X.slot[gi];
    if (src) {
      // the binned gather's output for this stage (kept in registers; only
      // stage 1's and stage 7's fields are stored, as ls_keep_field keeps them)
      const double2 r = L.X.rout[sl];
      const double x = sx[lt], y = sy[lt], rho = hypot(x, y);
      const double qx = rho > 0.0 ? r.x * (x / rho) : 0.0, qy = rho > 0.0 ? r.x * (y / rho) : 0.0;
      geometry_field_point(L.T.G, x, y, sz[lt], qx, qy, r.y, fe[0], fe[1], fe[2]);
      if (st == 0) {
        for (int c = 0; c < 3; ++c) LKB(54 + c) = fe[c];
        L.fvalid[gi] = 1;
      } else if (st == 6) {
        for (int c = 0; c < 3; ++c) LKB(57 + c) = fe[c];
      }
      const int key = L.X.rkey[sl];
      L.X.key[gi] = key;
      hint = bin_tile(L.X, key);
    } else {
      for (int c = 0; c < 3; ++c) fe[c] = LKB(51 + c);
    }
  }
  __syncthreads();
  if (!LAST && !mine) return;
  if (mine) {
    double Fc[3], Fi[3];
    const int b = gs * W;
    pair_forces_p<PR>(L.T.G, tid, n, sx + b, sy + b, sz + b, sq + b, act + b, L.T.a2, Fc, Fi);
    if (L.imgF)
      for (int c = 0; c < 3; ++c) Fi[c] = L.imgF[3 * gi + c];
    double kst[3];  // this stage's k (velocity part): the acceleration
    for (int c = 0; c < 3; ++c) {
      const double ac = n == 1 ? 0.0 : Fc[c] / mass;
      const double a = L.T.q_over_m * fe[c] + ac + Fi[c] / mass;
      kst[c] = active ? a : 0.0;
      LKB(st * 6 + 3 + c) = kst[c];
    }
    if (st < 6) {
      // The position part of stage j's k is stage j's velocity, not stored:
      // formed again from the stored accelerations with the same products in
      // the same order as when stage j was made (so bit for bit the same).
      double ps[3];  // the next stage's position
      for (int c = 0; c < 3; ++c) {
        const double p0c = p0[c], v0c = v0[c];
        double kv[6], kp[6];
  #pragma unroll
        for (int j = 0; j < 6; ++j) {
          kv[j] = j < st ? kprev[j < 5 ? j : 4][c] : (j == st ? kst[c] : 0.0);
          double v = v0c;
  #pragma unroll
          for (int i = 0; i < j; ++i) v = v + (hr * c_A[j][i]) * kv[i];
          kp[j] = active ? v : 0.0;
        }
        if (st < 5) {
          const int ns = st + 1;
          double pc = p0c;
  #pragma unroll
          for (int j = 0; j < 5; ++j) {
            if (j < ns) pc = pc + (hr * c_A[ns][j]) * kp[j];
          }
          ps[c] = pc;
          LKB(48 + c) = pc;
        } else {
          double sp = c_B[0] * kp[0], sv = c_B[0] * kv[0];
  #pragma unroll
          for (int k = 1; k < 6; ++k) {
            sp = sp + c_B[k] * kp[k];
            sv = sv + c_B[k] * kv[k];
          }
          const double pn = p0c + hr * sp;
          const double vn = v0c + hr * sv;
          LKB(42 + c) = pn;
          LKB(45 + c) = vn;
          LKB(48 + c) = pn;
          ps[c] = pn;
        }
      }
      if (binned) ls_bin_pack(L, gi, st + 1, ps[0], ps[1], ps[2], sl, active, hint);
    }
  }
  // every lane of a running group, together: the end's reductions are
  // shuffles over the group's lanes
  if (LAST && run) ls_end_group<TG, GPB>(L, g, flags, tid, sx);
}

// Error norm, accept/reject, commit, birth reset, kill, screens (reference
// batched_rk45.py:250-317, 372-399), then the group's next attempt.
// The end of group g's attempt, on the group's lanes (tid: the lane's
// member index). With GPB > 1 every lane of a running group must be here
// (its reductions are 16-lane shuffles); with GPB = 1 the whole block.
template <int TG, int GPB>
__device__ void ls_end_group(const LockParams& L, int g, int flags, int tid, double* red) {
  const bool live = flags & LS_RUNNING;
  const int64_t off = live ? L.T.offsets[g] : 0;
  const int n = live ? (int)(L.T.offsets[g + 1] - off) : 1;
  const bool mine = live && tid < n;
  const int64_t gi = off + tid;
  const double hr = live ? L.g_hr[g] : 0.0;
  double lp = 0.0, lv = 0.0;
  if (mine) {
    const bool active = L.active[gi];
    for (int c = 0; c < 3; ++c) {
      // the position parts of the k: the stage velocities, formed again as
      // the accel kernel forms them (stage 6's is the new velocity)
      const double v0 = L.T.vel[3 * gi + c], vn = LKB(45 + c);
      double kv[7], kp[7];
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        kv[k] = LKB(k * 6 + 3 + c);
        double v = v0;
        if (k < 6) {
#pragma unroll
          for (int i = 0; i < k; ++i) v = v + (hr * c_A[k][i]) * kv[i];
        } else {
          v = vn;
        }
        kp[k] = active ? v : 0.0;
      }
      double ep = c_E[0] * kp[0], ev = c_E[0] * kv[0];
      for (int k = 1; k < 7; ++k) {
        ep = ep + c_E[k] * kp[k];
        ev = ev + c_E[k] * kv[k];
      }
      ep = hr * ep;
      ev = hr * ev;
      const double sp = L.T.atol + L.T.rtol * fmax(fabs(L.T.pos[3 * gi + c]), fabs(LKB(42 + c)));
      const double sv = kLight * L.T.atol + L.T.rtol * fmax(fabs(L.T.vel[3 * gi + c]), fabs(LKB(45 + c)));
      const double a = ep / sp, b = ev / sv;
      lp += a * a;
      lv += b * b;
    }
  }
  const double sq = group_sum<TG, GPB>(lp, red) + group_sum<TG, GPB>(lv, red);
  if (GPB > 1 && !live) return;  // lanes of a finished group: no barrier follows in the packed form
  double err = sqrt(sq / (6.0 * n));
  if (err == 0.0) err = 2.2250738585072014e-308;
  const bool acc = err <= 1.0;
  double fac = L.T.safety * pow(err, -1.0 / 5);
  fac = fmin(fmax(fac, L.T.fmin_), L.T.fmax_);
  const int64_t batch = L.T.attempts[g];
  double t = L.g_t[g];
  double h = hr * fac;
  bool running = true;
  const int optr = outputs_reached(L.T, t);
  const bool oexh = optr >= L.T.n_out;
  const double nout = oexh ? L.T.t_max : L.T.t_out[optr];
  if (acc) {
    const double t_old = t;
    t = t + hr;
    const double tol = 1e-9 * fmax(L.T.t_max, 1e-30);
    if (!(flags & LS_EXHAUSTED) && t >= L.g_next[g] - tol) h = (L.T.t_max - t) * L.T.init_frac;
    if (mine) {
      double x0[3], v0[3], x1[3], v1[3];
      for (int c = 0; c < 3; ++c) {
        x0[c] = L.T.pos[3 * gi + c];
        v0[c] = L.T.vel[3 * gi + c];
        x1[c] = LKB(42 + c);
        v1[c] = LKB(45 + c);
        L.T.pos[3 * gi + c] = x1[c];
        L.T.vel[3 * gi + c] = v1[c];
      }
      if (L.active[gi]) {  // stage 7 sat at the new state
        for (int c = 0; c < 3; ++c) LKB(54 + c) = LKB(57 + c);
        L.fvalid[gi] = 1;
      }
      if (L.active[gi] && kill_point(L.T.G, x1[0], x1[1], x1[2])) L.T.t_death[gi] = t;
      screens_particle_body(L.T.SP, L.T.events, L.T.ev_cap, L.T.ev_count, g, tid, batch, t_old, t, x0, v0, x1, v1);
      if (!oexh && t >= nout - tol && L.T.t_birth[gi] <= t) record_snapshot(L.T, optr, g, tid, batch, t, x1, v1);
    }
    running = t < L.T.t_max * (1 - 1e-14);
  }
  const bool stalled = step_stalled(acc, h, t);
  if (stalled) running = false;
  if (GPB == 1) __syncthreads();
  if (tid == 0) {
    L.T.attempts[g] = batch + 1;
    if (acc) L.T.accepted[g] += 1;
    L.T.margin[g] = fmin(L.T.margin[g], fabs(err - 1.0));
    L.T.active_stages[g] += 7 * (int64_t)L.g_nact[g];
    L.g_t[g] = t;
    L.g_h[g] = h;
    if (L.T.t_end) L.T.t_end[g] = t;
    if (stalled) L.T.status[g] = SPT_STATUS_STALLED;
    if (!running) {
      L.g_flags[g] = 0;
      atomicAdd(L.done, 1u);
    }
  }
  if (!running) return;
  ls_begin<TG, GPB>(L, g, n, tid, t, h, batch + 1, red);
}

template <int TG, int GPB = 1>
__global__ __launch_bounds__(TG) void ls_end_kernel(LockParams L) {
  constexpr int W = TG / GPB;  // lanes per group
  __shared__ double red[TG];
  const int gs = threadIdx.x / W;
  const int g = blockIdx.x * GPB + gs;
  const int flags = g < L.T.n_groups ? L.g_flags[g] : 0;
  const bool live = flags & LS_RUNNING;
  if (GPB == 1 && !live) return;  // block-uniform
  if (GPB > 1 && !__syncthreads_or(live)) return;
  ls_end_group<TG, GPB>(L, g, flags, threadIdx.x - gs * W, red);
}
#undef LKB

// In-group pair forces: one block per group, threads over its particles.
template <int NMAX, int PR>
__global__ __launch_bounds__(NMAX) void pair_forces_kernel(spt_geometry G, const double* pos, const double* charge,
                                                           const uint8_t* active, const int64_t* offsets, double a2,
                                                           double* Fc_out, double* Fi_out) {
  __shared__ double sx[NMAX], sy[NMAX], sz[NMAX], sq[NMAX];
  __shared__ int act[NMAX];
  const int g = blockIdx.x;
  const int64_t off = offsets[g];
  const int n = (int)(offsets[g + 1] - off);
  const int i = threadIdx.x;
  if (i < n) {
    sx[i] = pos[3 * (off + i)];
    sy[i] = pos[3 * (off + i) + 1];
    sz[i] = pos[3 * (off + i) + 2];
    sq[i] = charge[off + i];
    act[i] = active[off + i] ? 1 : 0;
  }
  __syncthreads();
  if (i < n) {
    double Fc[3], Fi[3];
    pair_forces_p<PR>(G, i, n, sx, sy, sz, sq, act, a2, Fc, Fi);
    for (int c = 0; c < 3; ++c) {
      if (Fc_out) Fc_out[3 * (off + i) + c] = Fc[c];
      if (Fi_out) Fi_out[3 * (off + i) + c] = Fi[c];
    }
  }
}

__global__ void kill_kernel(spt_geometry G, const double* pos, int64_t n, uint8_t* out) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  out[k] = kill_point(G, pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]) ? 1 : 0;
}

// ------------------------------------------------------- BEM assembly
// Entry (i, j) of the hat-function collocation operator, split as
// lo[i, j] (hat of node j over segment j) and hi[i, j] (hat of node j+1
// over segment j): adaptive bisection while length > refine_ratio *
// distance(node i, midpoint), depth <= max_depth, Gauss-4 leaves
// (reference bem/axisymmetric.py:275-302). Depth-first with an explicit
// stack; leaf order matches the reference's left-then-right recursion.
__global__ void assemble_kernel(const double* prof, int n, double refine_ratio, int max_depth, double* lo_out,
                                double* hi_out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)n * (n - 1);
  if (k >= total) return;
  const int i = (int)(k / (n - 1)), j = (int)(k - (int64_t)i * (n - 1));
  const double fr = prof[2 * i], fz = prof[2 * i + 1];
  double st[48][6];
  int sd[48];
  int top = 0;
  st[0][0] = prof[2 * j];
  st[0][1] = prof[2 * j + 1];
  st[0][2] = prof[2 * j + 2];
  st[0][3] = prof[2 * j + 3];
  st[0][4] = 0.0;
  st[0][5] = 1.0;
  sd[0] = 0;
  top = 1;
  double lo = 0.0, hi = 0.0;
  while (top > 0) {
    --top;
    const double a0r = st[top][0], a0z = st[top][1], a1r = st[top][2], a1z = st[top][3];
    const double u0 = st[top][4], u1 = st[top][5];
    const int depth = sd[top];
    const double mr = 0.5 * (a0r + a1r), mz = 0.5 * (a0z + a1z);
    const double dr = a1r - a0r, dz = a1z - a0z;
    const double length = sqrt(dr * dr + dz * dz);
    const double er = fr - mr, ez = fz - mz;
    const double dist = sqrt(er * er + ez * ez);
    if (depth < max_depth && length > refine_ratio * dist && top + 2 <= 48) {
      const double um = 0.5 * (u0 + u1);
      // push right then left so the left half is integrated first
      st[top][0] = mr; st[top][1] = mz; st[top][2] = a1r; st[top][3] = a1z; st[top][4] = um; st[top][5] = u1;
      sd[top] = depth + 1;
      ++top;
      st[top][0] = a0r; st[top][1] = a0z; st[top][2] = mr; st[top][3] = mz; st[top][4] = u0; st[top][5] = um;
      sd[top] = depth + 1;
      ++top;
      continue;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double t = 0.5 * (c_gx4[q] + 1.0);
      const double rs = a0r + t * (a1r - a0r);
      const double zs = a0z + t * (a1z - a0z);
      const double hs = u0 + t * (u1 - u0);
      const double ls = (1.0 - u0) + t * ((1.0 - u1) - (1.0 - u0));
      const double gq = 2.0 * kPi * rs * ring_potential(fr, fz - zs, rs) * (length * 0.5 * c_gw4[q]);
      lo += ls * gq;
      hi += hs * gq;
    }
  }
  lo_out[k] = lo;
  hi_out[k] = hi;
}

// A[i, j] = hi[i, j-1] + lo[i, j], column-major for LAPACK-style solvers,
// and the Dirichlet right-hand side E_gun * z_i.
__global__ void assemble_combine_kernel(const double* prof, int n, const double* lo, const double* hi, double E_gun,
                                        double* A_cm, double* b) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)n * n) return;
  const int j = (int)(k / n), i = (int)(k - (int64_t)j * n);
  double a = 0.0;
  if (j > 0) a = a + hi[(int64_t)i * (n - 1) + (j - 1)];
  if (j < n - 1) a = a + lo[(int64_t)i * (n - 1) + j];
  A_cm[k] = a;
  if (j == 0 && b) b[i] = E_gun * prof[2 * i + 1];
}

// Perturbation potential at points (reference bem/axisymmetric.py:445-462):
// per point, the segments in order, each by adaptive bisection (split
// while length > refine_ratio * distance to the midpoint, depth <=
// max_depth) with Gauss-4 leaves and sigma linear along the segment
// (reference bem/axisymmetric.py:275-302). Depth-first, left half first.
// The part of the fixed-rule perturbation field a tile tabulates: at point i
// (rho, z), the sum over every segment outside [win[i].x, win[i].y) of its
// sub-panel Gauss nodes (reference bem/axisymmetric.py:560-564; these
// segments lie at least a panel length from the point, where the reference
// takes exactly this rule). One wavefront per point, lanes over the nodes,
// a fixed butterfly over the lanes.
__global__ __launch_bounds__(256) void remainder_kernel(BemView B, const double* pts, const int2* win, int64_t n,
                                                         double* out) {
  __shared__ double2 logt[SPT_LOGT_N];
  fill_log_table<256>(logt);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const double re = pts[2 * i], ze = pts[2 * i + 1], re2 = re * re;
  const int2 w = win[i];
  RingAcc acc;
  acc.zero();
  const int nq = B.nq, nn = B.n_seg * nq;
  for (int q = lane; q < nn; q += 64) {
    const int j = q / nq;
    if (j >= w.x && j < w.y) continue;
    const double4 v = B.nodes[q];
    ring_accum(re, re2, ze - v.y, v.x, v.z, acc, logt);
  }
  for (int o = 32; o > 0; o >>= 1) {
    acc.er_main += __shfl_xor(acc.er_main, o);
    acc.ez_main += __shfl_xor(acc.ez_main, o);
    acc.er_ax += __shfl_xor(acc.er_ax, o);
    acc.ez_ax += __shfl_xor(acc.ez_ax, o);
  }
  if (lane == 0) {
    double er, ez;
    ring_finish(acc, re, er, ez);
    out[2 * i] = er;
    out[2 * i + 1] = ez;
  }
}

__global__ void potential_kernel(const double* prof, const double* sigma, int n, double refine_ratio, int max_depth,
                                 const double* pos, int64_t np, double* phi) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const double fr = hypot(pos[3 * k], pos[3 * k + 1]), fz = pos[3 * k + 2];
  double st[48][6];
  int sd[48];
  double total = 0.0;
  for (int j = 0; j < n - 1; ++j) {
    st[0][0] = prof[2 * j];
    st[0][1] = prof[2 * j + 1];
    st[0][2] = prof[2 * j + 2];
    st[0][3] = prof[2 * j + 3];
    st[0][4] = sigma[j];
    st[0][5] = sigma[j + 1];
    sd[0] = 0;
    int top = 1;
    double seg = 0.0;
    while (top > 0) {
      --top;
      const double a0r = st[top][0], a0z = st[top][1], a1r = st[top][2], a1z = st[top][3];
      const double s0 = st[top][4], s1 = st[top][5];
      const int depth = sd[top];
      const double mr = 0.5 * (a0r + a1r), mz = 0.5 * (a0z + a1z);
      const double dr = a1r - a0r, dz = a1z - a0z;
      const double length = sqrt(dr * dr + dz * dz);
      const double er = fr - mr, ez = fz - mz;
      const double dist = sqrt(er * er + ez * ez);
      if (depth < max_depth && length > refine_ratio * dist && top + 2 <= 48) {
        const double sm = 0.5 * (s0 + s1);
        st[top][0] = mr; st[top][1] = mz; st[top][2] = a1r; st[top][3] = a1z; st[top][4] = sm; st[top][5] = s1;
        sd[top] = depth + 1;
        ++top;
        st[top][0] = a0r; st[top][1] = a0z; st[top][2] = mr; st[top][3] = mz; st[top][4] = s0; st[top][5] = sm;
        sd[top] = depth + 1;
        ++top;
        continue;
      }
      double leaf = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double t = 0.5 * (c_gx4[q] + 1.0);
        const double rs = a0r + t * (a1r - a0r);
        const double zs = a0z + t * (a1z - a0z);
        const double ss = s0 + t * (s1 - s0);
        leaf += ss * 2.0 * kPi * rs * ring_potential(fr, fz - zs, rs) * (length * 0.5 * c_gw4[q]);
      }
      seg += leaf;
    }
    total += seg;
  }
  phi[k] = total;
}

// ================================================================ host side
constexpr int FIELD_NMAX_PUBLIC = 64;  // largest points-per-block of the field gather

// Scratch of spt_trace_noalloc (the persistent trace without allocation):
// the work-queue counter in the first 256 bytes, then the stage buffer of
// kSlots doubles per particle. One copy per device, held until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// returns for the context's device (spt_ctx::static_scratch).
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
constexpr size_t kStaticKbufOffset = 256 / sizeof(double);
constexpr size_t kStaticKbufElems = kStaticScratchBytes / sizeof(double) - kStaticKbufOffset;
static __device__ __align__(256) double scratch[kStaticScratchBytes / sizeof(double)];

struct spt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_cu = 0;
  std::string err;
  // BEM tables
  int bem_nodes = 0, bem_sub = 0;
  double* d_prof = nullptr;
  double* d_sigma = nullptr;
  double* d_segs = nullptr;
  double4* d_nodes = nullptr;
  double4* d_coarse = nullptr;      // coarse far-rule nodes (kCoarsePer per segment)
  double far_k4 = 0.0, far_k2 = 0.0;  // spt_bem_far_rule; 0 = the reference's fixed rule
  double4* d_mp = nullptr;            // spt_bem_multipole table (kMultipoleMaxTerms entries)
  int mp_n = 0;
  double mp_rref = 0.0, mp_rmin = 0.0;
  // spt_bem_clusters: one allocation holding the cluster table, the
  // moments and the recurrence coefficients; the segment ranges and the
  // per-segment cluster index in another
  double* d_cl = nullptr;
  int* d_cl_idx = nullptr;
  int cl_n = 0, cl_terms = 0;
  double cl_tol = 0.0;
  double near_w = kRefNearWidth;      // spt_bem_near_width
  // spt_bem_tiles: quadtree, windows and coefficients (one allocation
  // each), used by far-rule gathers while tl_on
  int* d_tl_tree = nullptr;
  int2* d_tl_win = nullptr;
  double* d_tl_coef = nullptr;
  double4* d_tl_cell = nullptr;
  int tl_n = 0;
  int tl_nwin = 0;  // tiles with a window (numbered first)
  bool tl_on = false;
  double tl_x0 = 0.0, tl_z0 = 0.0, tl_size = 0.0;
  // scratch
  double* d_kbuf = nullptr;
  size_t kbuf_elems = 0;
  int* d_queue = nullptr;
  double* static_scratch = nullptr;  // this device's `scratch` (hipGetSymbolAddress), or nullptr
  std::string static_err;
  rocblas_handle blas = nullptr;
  // lockstep scratch
  hipStream_t own = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  int32_t* d_pgroup = nullptr;
  uint8_t* d_active = nullptr;
  size_t part_cap = 0;
  double* d_gdbl = nullptr;   // 5 x group doubles
  int32_t* d_gflags = nullptr;
  size_t group_cap = 0;
  unsigned int* d_done = nullptr;
  unsigned int* h_done = nullptr;
  uint8_t* d_sur = nullptr;
  uint8_t* d_fvalid = nullptr;
  // binned gather of far-rule lockstep traces (spt_set_lockstep_binning)
  int bin_mode = 1;  // 0 off, 1 on, 2 append path only (a test of slot-order independence)
  unsigned* d_bin_ctl = nullptr;  // hist, cursor (kBinMaxKeys each), total
  int32_t* d_bin_key = nullptr;
  int32_t* d_bin_slot = nullptr;
  uint8_t* d_bin_src = nullptr;
  double2* d_bin_rec = nullptr;   // rin, rout (n double2 each), then rtag, rkey (n int32 each)
  int32_t* d_bin_glist = nullptr; // the compact list of gather records (n)
  size_t bin_cap = 0;
  // surrogate used by spt_trace
  bool have_sur = false;
  spt_mlp_cfg sur_cfg;
  float* d_sur_p = nullptr;
  __bf16* d_sur_split = nullptr;
  double* d_sur_frag = nullptr;
  // MLP scratch: split weights for spt_mlp_infer, training buffers
  __bf16* d_mlp_split = nullptr;
  double* d_mlp_frag = nullptr;
  float* d_train = nullptr;
  size_t train_floats = 0;
  int train_batch = 0;
  // field-gather launch shape chosen by a tuner (0 = automatic)
  int tune_threads = 0, tune_pb = 0;
  // image-charge BEM table and per-call scratch
  int img_P = 0, img_modes = 0, img_quad = 0;
  double img_floor = 0.0, img_dlo = 0.0, img_dhi = 0.0;
  double* d_img_prof = nullptr;
  double* d_img_segs = nullptr;
  double* d_img_quadv = nullptr;
  double* d_img_Ainv = nullptr;
  double* d_img_rhs = nullptr;
  double* d_img_sigma = nullptr;
  size_t img_groups_cap = 0;
  double* d_img_mirror = nullptr;
  double* d_img_F = nullptr;
  size_t img_part_cap = 0;
};

// Every entry point starts here: make the context's device current and
// clear any error another library left behind on this thread, so the
// launch checks below see only this call's errors.
static hipError_t spt_begin(spt_ctx* c) {
  const hipError_t e = hipSetDevice(c->device);
  (void)hipGetLastError();
  return e;
}

static int fail(spt_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(ctx, SPT_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
  } while (0)

static BemView bem_view(const spt_ctx* c) {
  BemView B;
  B.n_seg = c->bem_nodes - 1;
  B.nq = c->bem_sub * 4;
  B.nodes = c->d_nodes;
  B.segs = c->d_segs;
  B.sigma = c->d_sigma;
  B.coarse = c->far_k4 > 0.0 ? c->d_coarse : nullptr;
  B.k4 = c->far_k4;
  B.k2 = c->far_k2;
  B.mp = c->mp_n > 0 ? c->d_mp : nullptr;
  B.mp_n = c->mp_n;
  B.mp_rref = c->mp_rref;
  B.mp_rmin2 = c->mp_rmin * c->mp_rmin;
  const bool cl = c->far_k4 > 0.0 && c->cl_n > 0;
  B.cl = cl ? reinterpret_cast<const double4*>(c->d_cl) : nullptr;
  B.cl_mom = cl ? c->d_cl + 4 * (size_t)c->cl_n : nullptr;
  B.cl_coef = cl ? reinterpret_cast<const double2*>(c->d_cl + 4 * (size_t)c->cl_n + (size_t)c->cl_n * c->cl_terms)
                 : nullptr;
  B.cl_seg = cl ? c->d_cl_idx : nullptr;
  B.seg_cl = cl ? c->d_cl_idx + c->cl_n + 1 : nullptr;
  B.cl_n = cl ? c->cl_n : 0;
  B.cl_terms = c->cl_terms;
  B.cl_tol = c->cl_tol;
  B.near_w = c->far_k4 > 0.0 ? c->near_w : kRefNearWidth;
  const bool tl = c->far_k4 > 0.0 && c->tl_on && c->tl_n > 0;
  B.tl_tree = tl ? c->d_tl_tree : nullptr;
  B.tl_win = tl ? c->d_tl_win : nullptr;
  B.tl_coef = tl ? c->d_tl_coef : nullptr;
  B.tl_cell = tl ? c->d_tl_cell : nullptr;
  B.tl_n = tl ? c->tl_n : 0;
  B.tl_nwin = tl ? c->tl_nwin : 0;
  B.tl_x0 = c->tl_x0;
  B.tl_z0 = c->tl_z0;
  B.tl_size = c->tl_size;
  return B;
}

static void free_tiles(spt_ctx* c) {
  if (c->d_tl_tree) (void)hipFree(c->d_tl_tree);
  if (c->d_tl_win) (void)hipFree(c->d_tl_win);
  if (c->d_tl_coef) (void)hipFree(c->d_tl_coef);
  if (c->d_tl_cell) (void)hipFree(c->d_tl_cell);
  c->d_tl_tree = nullptr;
  c->d_tl_win = nullptr;
  c->d_tl_coef = nullptr;
  c->d_tl_cell = nullptr;
  c->tl_n = 0;
  c->tl_on = false;
}

static void free_bem(spt_ctx* c) {
  if (c->d_prof) (void)hipFree(c->d_prof);
  if (c->d_sigma) (void)hipFree(c->d_sigma);
  if (c->d_segs) (void)hipFree(c->d_segs);
  if (c->d_nodes) (void)hipFree(c->d_nodes);
  if (c->d_coarse) (void)hipFree(c->d_coarse);
  c->d_prof = c->d_sigma = c->d_segs = nullptr;
  c->d_nodes = c->d_coarse = nullptr;
  c->mp_n = 0;  // the expansions and tiles belong to the uploaded charge
  c->cl_n = 0;
  free_tiles(c);
  c->bem_nodes = 0;
}

extern "C" {

int spt_create(int device, spt_ctx** out) {
  if (!out) return SPT_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return SPT_EINVAL;
  spt_ctx* c = new spt_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return SPT_EHIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete c;
    return SPT_EHIP;
  }
  c->n_cu = prop.multiProcessorCount;
  if (hipMalloc(&c->d_queue, 64) != hipSuccess) {
    delete c;
    return SPT_ENOMEM;
  }
  // the device address of this device's copy of `scratch` (the host-side
  // name is only a stand-in): everything that passes it to a kernel uses it
  const hipError_t se = hipGetSymbolAddress((void**)&c->static_scratch, HIP_SYMBOL(scratch));
  if (se != hipSuccess) {
    c->static_scratch = nullptr;
    c->static_err = hipGetErrorString(se);
    (void)hipGetLastError();
  }
  *out = c;
  return SPT_OK;
}

void spt_destroy(spt_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  free_bem(c);
  if (c->d_mp) (void)hipFree(c->d_mp);
  if (c->d_cl) (void)hipFree(c->d_cl);
  if (c->d_cl_idx) (void)hipFree(c->d_cl_idx);
  if (c->d_kbuf) (void)hipFree(c->d_kbuf);
  if (c->d_queue) (void)hipFree(c->d_queue);
  if (c->d_pgroup) (void)hipFree(c->d_pgroup);
  if (c->d_active) (void)hipFree(c->d_active);
  if (c->d_gdbl) (void)hipFree(c->d_gdbl);
  if (c->d_gflags) (void)hipFree(c->d_gflags);
  if (c->d_done) (void)hipFree(c->d_done);
  if (c->h_done) (void)hipHostFree(c->h_done);
  if (c->d_sur) (void)hipFree(c->d_sur);
  if (c->d_fvalid) (void)hipFree(c->d_fvalid);
  if (c->d_bin_ctl) (void)hipFree(c->d_bin_ctl);
  if (c->d_bin_key) (void)hipFree(c->d_bin_key);
  if (c->d_bin_slot) (void)hipFree(c->d_bin_slot);
  if (c->d_bin_src) (void)hipFree(c->d_bin_src);
  if (c->d_bin_rec) (void)hipFree(c->d_bin_rec);
  if (c->d_bin_glist) (void)hipFree(c->d_bin_glist);
  if (c->d_sur_p) (void)hipFree(c->d_sur_p);
  if (c->d_sur_split) (void)hipFree(c->d_sur_split);
  if (c->d_sur_frag) (void)hipFree(c->d_sur_frag);
  if (c->d_mlp_frag) (void)hipFree(c->d_mlp_frag);
  if (c->d_mlp_split) (void)hipFree(c->d_mlp_split);
  if (c->d_train) (void)hipFree(c->d_train);
  for (double* p : {c->d_img_prof, c->d_img_segs, c->d_img_quadv, c->d_img_Ainv, c->d_img_rhs, c->d_img_sigma,
                    c->d_img_mirror, c->d_img_F})
    if (p) (void)hipFree(p);
  if (c->ev_in) (void)hipEventDestroy(c->ev_in);
  if (c->ev_out) (void)hipEventDestroy(c->ev_out);
  if (c->own) (void)hipStreamDestroy(c->own);
  if (c->blas) rocblas_destroy_handle(c->blas);
  delete c;
}

const char* spt_last_error(const spt_ctx* c) { return c ? c->err.c_str() : "null context"; }

int spt_set_stream(spt_ctx* c, void* s) {
  if (!c) return SPT_EINVAL;
  c->stream = (hipStream_t)s;
  if (c->blas) rocblas_set_stream(c->blas, c->stream);
  return SPT_OK;
}

int spt_sync(spt_ctx* c) {
  if (!c) return SPT_EINVAL;
  HIPCHK(c, spt_begin(c));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPT_OK;
}

int spt_set_field_shape(spt_ctx* c, int32_t threads, int32_t points_per_block) {
  if (!c) return SPT_EINVAL;
  if (threads != 0 && threads != 64 && threads != 128 && threads != 192 && threads != 256)
    return fail(c, SPT_EINVAL, "threads must be 0, 64, 128, 192 or 256");
  if (points_per_block < 0 || points_per_block > FIELD_NMAX_PUBLIC ||
      (points_per_block & (points_per_block - 1)) != 0)
    return fail(c, SPT_EINVAL, "points_per_block must be 0 or a power of two <= 64");
  c->tune_threads = threads;
  c->tune_pb = points_per_block;
  return SPT_OK;
}

int spt_set_lockstep_binning(spt_ctx* c, int32_t on) {
  if (!c) return SPT_EINVAL;
  if (on < 0 || on > 2) return fail(c, SPT_EINVAL, "spt_set_lockstep_binning: on must be 0, 1 or 2");
  c->bin_mode = on;
  return SPT_OK;
}

int spt_device_cus(spt_ctx* c, int* n) {
  if (!c || !n) return SPT_EINVAL;
  *n = c->n_cu;
  return SPT_OK;
}

int spt_bem_upload(spt_ctx* c, const double* prof, int32_t n_nodes, const double* sigma, int32_t n_sub) {
  if (!c) return SPT_EINVAL;
  if (!prof || !sigma || n_nodes < 2 || n_sub < 1 || n_sub > 64)
    return fail(c, SPT_EINVAL, "spt_bem_upload: need n_nodes >= 2 and 1 <= n_subdiv <= 64");
  HIPCHK(c, spt_begin(c));
  free_bem(c);
  const int n_seg = n_nodes - 1;
  const size_t nq = (size_t)n_seg * n_sub * 4;
  HIPCHK(c, hipMalloc(&c->d_prof, sizeof(double) * 2 * n_nodes));
  HIPCHK(c, hipMalloc(&c->d_sigma, sizeof(double) * n_nodes));
  HIPCHK(c, hipMalloc(&c->d_segs, sizeof(double) * SEG_STRIDE * n_seg));
  HIPCHK(c, hipMalloc(&c->d_nodes, sizeof(double4) * nq));
  HIPCHK(c, hipMalloc(&c->d_coarse, sizeof(double4) * kCoarsePer * n_seg));
  HIPCHK(c, hipMemcpyAsync(c->d_prof, prof, sizeof(double) * 2 * n_nodes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_sigma, sigma, sizeof(double) * n_nodes, hipMemcpyHostToDevice, c->stream));
  seg_table_kernel<<<(n_seg + 255) / 256, 256, 0, c->stream>>>(c->d_prof, n_seg, c->d_segs);
  node_table_kernel<<<(unsigned)((nq + 255) / 256), 256, 0, c->stream>>>(c->d_prof, c->d_sigma, n_seg, n_sub,
                                                                          c->d_nodes);
  coarse_table_kernel<<<(n_seg * kCoarsePer + 255) / 256, 256, 0, c->stream>>>(c->d_prof, c->d_sigma, n_seg,
                                                                              c->d_coarse);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->bem_nodes = n_nodes;
  c->bem_sub = n_sub;
  return SPT_OK;
}

int spt_bem_far_rule(spt_ctx* c, double k4, double k2) {
  if (!c) return SPT_EINVAL;
  if (k4 == 0.0) {
    c->far_k4 = c->far_k2 = 0.0;
    return SPT_OK;
  }
  if (!(k4 >= 1.0 && k2 >= k4)) return fail(c, SPT_EINVAL, "spt_bem_far_rule: need k4 = 0, or 1 <= k4 <= k2");
  c->far_k4 = k4;
  c->far_k2 = k2;
  return SPT_OK;
}

int spt_bem_multipole(spt_ctx* c, const double* moments, int32_t n_terms, double r_ref, double r_min) {
  if (!c) return SPT_EINVAL;
  if (n_terms == 0) {
    c->mp_n = 0;
    return SPT_OK;
  }
  if (!moments || n_terms < 1 || n_terms > kMultipoleMaxTerms || !(r_ref > 0.0) || !(r_min > r_ref))
    return fail(c, SPT_EINVAL, "spt_bem_multipole: need 1 <= n_terms <= 1024 and 0 < r_ref < r_min (or n_terms = 0)");
  if (c->bem_nodes == 0) return fail(c, SPT_ENOBEM, "spt_bem_multipole: no BEM table uploaded");
  HIPCHK(c, spt_begin(c));
  if (!c->d_mp) HIPCHK(c, hipMalloc(&c->d_mp, sizeof(double4) * kMultipoleMaxTerms));
  std::vector<double4> tab(n_terms);
  for (int l = 0; l < n_terms; ++l)
    tab[l] = make_double4(moments[l], (2.0 * l + 1.0) / (l + 1.0), (double)l / (l + 1.0), 0.0);
  HIPCHK(c, hipMemcpyAsync(c->d_mp, tab.data(), sizeof(double4) * n_terms, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->mp_n = n_terms;
  c->mp_rref = r_ref;
  c->mp_rmin = r_min;
  return SPT_OK;
}

int spt_bem_clusters(spt_ctx* c, const int32_t* seg_begin, int32_t n_clusters, const double* z_c, const double* r_c,
                     const double* moments, int32_t n_terms, double kappa, double tol) {
  if (!c) return SPT_EINVAL;
  if (n_clusters == 0) {
    c->cl_n = 0;
    return SPT_OK;
  }
  if (c->bem_nodes == 0) return fail(c, SPT_ENOBEM, "spt_bem_clusters: no BEM table uploaded");
  const int n_seg = c->bem_nodes - 1;
  if (!seg_begin || !z_c || !r_c || !moments || n_clusters < 1 || n_clusters > kMaxClusters || n_terms < 1 ||
      n_terms > kClusterMaxTerms || !(kappa > 1.0) || !(tol > 0.0))
    return fail(c, SPT_EINVAL, "spt_bem_clusters: need 1 <= n_clusters <= 64, 1 <= n_terms <= 256, kappa > 1, tol > 0");
  if (seg_begin[0] != 0 || seg_begin[n_clusters] != n_seg)
    return fail(c, SPT_EINVAL, "spt_bem_clusters: seg_begin must run from 0 to the segment count");
  for (int k = 0; k < n_clusters; ++k)
    if (seg_begin[k + 1] <= seg_begin[k] || !(r_c[k] > 0.0))
      return fail(c, SPT_EINVAL, "spt_bem_clusters: clusters need increasing segment ranges and r_c > 0");
  HIPCHK(c, spt_begin(c));
  const size_t nd = 4 * (size_t)n_clusters + (size_t)n_clusters * n_terms + 2 * (size_t)n_terms;
  std::vector<double> tab(nd);
  for (int k = 0; k < n_clusters; ++k) {
    tab[4 * k] = z_c[k];
    tab[4 * k + 1] = r_c[k];
    tab[4 * k + 2] = (kappa * r_c[k]) * (kappa * r_c[k]);
    tab[4 * k + 3] = 0.0;
  }
  std::copy(moments, moments + (size_t)n_clusters * n_terms, tab.begin() + 4 * n_clusters);
  for (int l = 0; l < n_terms; ++l) {
    tab[4 * n_clusters + (size_t)n_clusters * n_terms + 2 * l] = (2.0 * l + 1.0) / (l + 1.0);
    tab[4 * n_clusters + (size_t)n_clusters * n_terms + 2 * l + 1] = (double)l / (l + 1.0);
  }
  std::vector<int> idx(n_clusters + 1 + n_seg);
  for (int k = 0; k <= n_clusters; ++k) idx[k] = seg_begin[k];
  for (int k = 0; k < n_clusters; ++k)
    for (int j = seg_begin[k]; j < seg_begin[k + 1]; ++j) idx[n_clusters + 1 + j] = k;
  if (c->d_cl) (void)hipFree(c->d_cl);
  if (c->d_cl_idx) (void)hipFree(c->d_cl_idx);
  c->d_cl = nullptr;
  c->d_cl_idx = nullptr;
  c->cl_n = 0;
  HIPCHK(c, hipMalloc(&c->d_cl, sizeof(double) * nd));
  HIPCHK(c, hipMalloc(&c->d_cl_idx, sizeof(int) * idx.size()));
  HIPCHK(c, hipMemcpyAsync(c->d_cl, tab.data(), sizeof(double) * nd, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_cl_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->cl_n = n_clusters;
  c->cl_terms = n_terms;
  c->cl_tol = tol;
  return SPT_OK;
}

int spt_bem_near_width(spt_ctx* c, double width) {
  if (!c) return SPT_EINVAL;
  if (width == 0.0) width = kRefNearWidth;
  if (!(width > 0.0 && width <= 4.0)) return fail(c, SPT_EINVAL, "spt_bem_near_width: need 0 < width <= 4 (0: 0.5)");
  c->near_w = width;
  return SPT_OK;
}

int spt_bem_tiles(spt_ctx* c, const int32_t* tree, int32_t n_tree, const int32_t* win, const double* coef,
                  int32_t n_tiles, double x0, double z0, double size) {
  if (!c) return SPT_EINVAL;
  if (c->bem_nodes == 0) return fail(c, SPT_ENOBEM, "spt_bem_tiles: no BEM table uploaded");
  const int n_seg = c->bem_nodes - 1;
  if (!tree || !win || !coef || n_tree < 1 || n_tiles < 1 || n_tiles > kTileMaxTiles || !(size > 0.0) ||
      !std::isfinite(x0) || !std::isfinite(z0) || !std::isfinite(size))
    return fail(c, SPT_EINVAL, "spt_bem_tiles: need a tree, 1 <= n_tiles <= 6144 and a finite box of positive size");
  // The device walk trusts the tree: children in range and after their
  // parent (so every walk ends), tiles in range, windows within the profile.
  for (int k = 0; k < n_tree; ++k) {
    const int v = tree[k];
    if (v >= 0 ? (v <= k || v > n_tree - 4) : (v < -1 && -2 - v >= n_tiles))
      return fail(c, SPT_EINVAL, "spt_bem_tiles: tree node " + std::to_string(k) + " out of range");
  }
  for (int t = 0; t < n_tiles; ++t)
    if (win[2 * t] < 0 || win[2 * t + 1] < win[2 * t] || win[2 * t + 1] > n_seg)
      return fail(c, SPT_EINVAL, "spt_bem_tiles: window of tile " + std::to_string(t) + " out of range");
  int nwin = 0;  // tiles with a window come first
  while (nwin < n_tiles && win[2 * nwin + 1] > win[2 * nwin]) ++nwin;
  for (int t = nwin; t < n_tiles; ++t)
    if (win[2 * t + 1] > win[2 * t])
      return fail(c, SPT_EINVAL, "spt_bem_tiles: tiles with a window must come first (tile " + std::to_string(t) + ")");
  // every tile's cell with the device walk's arithmetic (tile_find: halve
  // the side, then add it to the low corner), so a cell lookup and a walk
  // agree to the bit; a tile reached twice or never is an error
  std::vector<double4> cell(n_tiles);
  std::vector<char> seen(n_tiles, 0);
  {
    struct Item { int k; double x0, z0, s; };
    std::vector<Item> stack{{0, x0, z0, size}};
    while (!stack.empty()) {
      const Item it = stack.back();
      stack.pop_back();
      const int v = tree[it.k];
      if (v <= -2) {
        const int t = -2 - v;
        if (seen[t]) return fail(c, SPT_EINVAL, "spt_bem_tiles: tile " + std::to_string(t) + " in two cells");
        seen[t] = 1;
        cell[t] = make_double4(it.x0, it.z0, it.s, 0.0);
      } else if (v >= 0) {
        double s = it.s;
        s *= 0.5;
        for (int q = 0; q < 4; ++q) {
          double cx0 = it.x0, cz0 = it.z0;
          if (q & 1) cx0 += s;
          if (q & 2) cz0 += s;
          stack.push_back({v + q, cx0, cz0, s});
        }
      }
    }
  }
  for (int t = 0; t < n_tiles; ++t)
    if (!seen[t]) return fail(c, SPT_EINVAL, "spt_bem_tiles: tile " + std::to_string(t) + " in no cell");
  HIPCHK(c, spt_begin(c));
  free_tiles(c);
  HIPCHK(c, hipMalloc(&c->d_tl_tree, sizeof(int) * n_tree));
  HIPCHK(c, hipMalloc(&c->d_tl_win, sizeof(int2) * n_tiles));
  HIPCHK(c, hipMalloc(&c->d_tl_coef, sizeof(double) * kTileCoef * (size_t)n_tiles));
  HIPCHK(c, hipMalloc(&c->d_tl_cell, sizeof(double4) * n_tiles));
  HIPCHK(c, hipMemcpyAsync(c->d_tl_cell, cell.data(), sizeof(double4) * n_tiles, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tl_tree, tree, sizeof(int) * n_tree, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tl_win, win, sizeof(int2) * n_tiles, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tl_coef, coef, sizeof(double) * kTileCoef * (size_t)n_tiles, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->tl_n = n_tiles;
  c->tl_nwin = nwin;
  c->tl_x0 = x0;
  c->tl_z0 = z0;
  c->tl_size = size;
  c->tl_on = true;
  return SPT_OK;
}

int spt_bem_tiles_enable(spt_ctx* c, int32_t on) {
  if (!c) return SPT_EINVAL;
  if (on && c->tl_n == 0) return fail(c, SPT_EINVAL, "spt_bem_tiles_enable: no tiles uploaded");
  c->tl_on = on != 0;
  return SPT_OK;
}

int spt_bem_remainder(spt_ctx* c, const double* pts, const int32_t* win, int64_t n, double* E) {
  if (!c) return SPT_EINVAL;
  if (!c->bem_nodes) return fail(c, SPT_ENOBEM, "no BEM table uploaded");
  if (n < 0 || (n > 0 && (!pts || !win || !E))) return fail(c, SPT_EINVAL, "spt_bem_remainder: bad arguments");
  if (n == 0) return SPT_OK;
  HIPCHK(c, spt_begin(c));
  remainder_kernel<<<(unsigned)((n + 3) / 4), 256, 0, c->stream>>>(bem_view(c), pts,
                                                                      reinterpret_cast<const int2*>(win), n, E);
  HIPCHK(c, hipGetLastError());
  return SPT_OK;
}

static int assemble_device(spt_ctx* c, const double* prof, int n, double rr, int md, double E_gun, double** A_cm,
                           double** b) {
  if (n < 2 || n > 20000) return fail(c, SPT_EINVAL, "assemble: need 2 <= n_nodes <= 20000");
  if (md < 0 || md > 40) return fail(c, SPT_EINVAL, "assemble: need 0 <= max_depth <= 40");
  double *d_prof = nullptr, *d_lo = nullptr, *d_hi = nullptr;
  const int64_t tot = (int64_t)n * (n - 1);
  HIPCHK(c, hipMalloc(&d_prof, sizeof(double) * 2 * n));
  HIPCHK(c, hipMalloc(&d_lo, sizeof(double) * tot));
  HIPCHK(c, hipMalloc(&d_hi, sizeof(double) * tot));
  HIPCHK(c, hipMalloc(A_cm, sizeof(double) * (size_t)n * n));
  HIPCHK(c, hipMalloc(b, sizeof(double) * n));
  HIPCHK(c, hipMemcpyAsync(d_prof, prof, sizeof(double) * 2 * n, hipMemcpyHostToDevice, c->stream));
  assemble_kernel<<<(unsigned)((tot + 127) / 128), 128, 0, c->stream>>>(d_prof, n, rr, md, d_lo, d_hi);
  assemble_combine_kernel<<<(unsigned)(((int64_t)n * n + 255) / 256), 256, 0, c->stream>>>(d_prof, n, d_lo, d_hi,
                                                                                           E_gun, *A_cm, *b);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(d_prof);
  (void)hipFree(d_lo);
  (void)hipFree(d_hi);
  return SPT_OK;
}

int spt_bem_assemble(spt_ctx* c, const double* prof, int32_t n, double rr, int32_t md, double* A_out) {
  if (!c || !prof || !A_out) return SPT_EINVAL;
  HIPCHK(c, spt_begin(c));
  double *A = nullptr, *b = nullptr;
  int rc = assemble_device(c, prof, n, rr, md, 0.0, &A, &b);
  if (rc) return rc;
  // column-major on device -> row-major host
  double* tmp = new double[(size_t)n * n];
  hipError_t e = hipMemcpy(tmp, A, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) A_out[(size_t)i * n + j] = tmp[(size_t)j * n + i];
  delete[] tmp;
  (void)hipFree(A);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(c, SPT_EHIP, hipGetErrorString(e));
  return SPT_OK;
}

int spt_bem_solve(spt_ctx* c, const double* prof, int32_t n, double E_gun, double rr, int32_t md, double* sigma_out) {
  if (!c || !prof || !sigma_out) return SPT_EINVAL;
  HIPCHK(c, spt_begin(c));
  double *A = nullptr, *b = nullptr;
  int rc = assemble_device(c, prof, n, rr, md, E_gun, &A, &b);
  if (rc) return rc;
  if (!c->blas) {
    if (rocblas_create_handle(&c->blas) != rocblas_status_success) return fail(c, SPT_EHIP, "rocblas_create_handle");
    rocblas_set_stream(c->blas, c->stream);
  }
  rocblas_int *ipiv = nullptr, *info = nullptr;
  HIPCHK(c, hipMalloc(&ipiv, sizeof(rocblas_int) * n));
  HIPCHK(c, hipMalloc(&info, sizeof(rocblas_int)));
  rocblas_status s1 = rocsolver_dgetrf(c->blas, n, n, A, n, ipiv, info);
  rocblas_status s2 = rocsolver_dgetrs(c->blas, rocblas_operation_none, n, 1, A, n, ipiv, b, n);
  rocblas_int h_info = 0;
  HIPCHK(c, hipMemcpyAsync(&h_info, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(sigma_out, b, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(ipiv);
  (void)hipFree(info);
  (void)hipFree(A);
  (void)hipFree(b);
  if (s1 != rocblas_status_success || s2 != rocblas_status_success) return fail(c, SPT_EHIP, "rocsolver LU failed");
  if (h_info != 0) return fail(c, SPT_EINVAL, "collocation matrix is singular");
  return SPT_OK;
}

#define FIELD_T 256
#define FIELD_NMAX 64

// Points per block for a field gather over n points: the largest power of
// two giving at least 48 blocks per compute unit (several rounds of three
// resident workgroups, so the last round's imbalance stays small; measured
// at 262,144 points: 16 per block 13.2 ms, 32 13.9 ms, 64 17.4 ms), at
// most FIELD_NMAX.
// Points per block for n points: 64, the gather's wave-major layout (node
// tables read by scalar loads), when that still gives every CU four rounds
// of three blocks; otherwise the largest power of two up to 32 with at
// least 48 blocks per CU.
static int field_points_per_block(const spt_ctx* c, int64_t n) {
  if (n >= (int64_t)FIELD_NMAX * 12 * c->n_cu) return FIELD_NMAX;
  const int64_t want = n / (48 * (int64_t)c->n_cu);
  int PB = 1;
  while (PB * 2 <= want && PB < FIELD_NMAX / 2) PB *= 2;
  return PB;
}

}  // extern "C"

// Field-gather launch shape for n points: points per block and block size
// (192 or 256 threads). At three resident waves per SIMD a compute unit
// holds four 3-wave or three 4-wave blocks; pick the shape whose block
// count fills whole rounds best (e.g. 1024 points on 256 CUs: 1024
// one-point blocks of 192 threads are exactly one round).
struct FieldShape {
  int T, PB;
};
static FieldShape field_shape(const spt_ctx* c, int64_t n, bool bem) {
  if (!bem) return {256, FIELD_NMAX};
  const int PB = c->tune_pb ? c->tune_pb : field_points_per_block(c, n);
  if (c->tune_threads) return {c->tune_threads, PB};
  const int64_t blocks = (n + PB - 1) / PB;
  double best = -1.0;
  FieldShape out{256, PB};
  for (int T : {192, 256}) {
    const int64_t cap = (int64_t)c->n_cu * (12 / (T / 64));
    const double rounds = (double)blocks / (double)cap;
    const double eff = rounds / std::ceil(rounds);
    if (eff > best + 1e-9) {
      best = eff;
      out = {T, PB};
    }
  }
  return out;
}

extern "C" {

int spt_bem_pert_field(spt_ctx* c, const double* pos, int64_t n, double* E, int32_t precision) {
  if (!c) return SPT_EINVAL;
  if (!c->bem_nodes) return fail(c, SPT_ENOBEM, "no BEM table uploaded");
  if (n < 0 || (n > 0 && (!pos || !E))) return fail(c, SPT_EINVAL, "spt_bem_pert_field: bad arguments");
  if (n == 0) return SPT_OK;
  HIPCHK(c, spt_begin(c));
  const FieldShape fs = field_shape(c, n, true);
  const unsigned nb = (unsigned)((n + fs.PB - 1) / fs.PB);
  if ((fs.T == 64 || fs.T == 128) && precision != SPT_PREC_MIXED) {
    // the binned lockstep gather's narrow shapes (one or two slices per
    // point at 64 points per block)
    if (fs.T == 64)
      bem_pert_kernel<64, FIELD_NMAX, SPT_PREC_FP64><<<nb, 64, 0, c->stream>>>(bem_view(c), pos, n, E, fs.PB);
    else
      bem_pert_kernel<128, FIELD_NMAX, SPT_PREC_FP64><<<nb, 128, 0, c->stream>>>(bem_view(c), pos, n, E, fs.PB);
  } else if (fs.T == 192) {
    if (precision == SPT_PREC_MIXED)
      bem_pert_kernel<192, FIELD_NMAX, SPT_PREC_MIXED><<<nb, 192, 0, c->stream>>>(bem_view(c), pos, n, E, fs.PB);
    else
      bem_pert_kernel<192, FIELD_NMAX, SPT_PREC_FP64><<<nb, 192, 0, c->stream>>>(bem_view(c), pos, n, E, fs.PB);
  } else {
    if (precision == SPT_PREC_MIXED)
      bem_pert_kernel<256, FIELD_NMAX, SPT_PREC_MIXED><<<nb, 256, 0, c->stream>>>(bem_view(c), pos, n, E, fs.PB);
    else
      bem_pert_kernel<256, FIELD_NMAX, SPT_PREC_FP64><<<nb, 256, 0, c->stream>>>(bem_view(c), pos, n, E, fs.PB);
  }
  HIPCHK(c, hipGetLastError());
  return SPT_OK;
}

static int check_geom(spt_ctx* c, const spt_geometry* g) {
  if (!g) return fail(c, SPT_EINVAL, "null geometry");
  if (g->kind < SPT_GEOM_FLAT || g->kind > SPT_GEOM_BEM_WELL) return fail(c, SPT_EINVAL, "unknown geometry kind");
  if (g->image < SPT_IMAGE_NONE || g->image > SPT_IMAGE_BEM_MIRROR) return fail(c, SPT_EINVAL, "unknown image model");
  if (is_image_bem(g->image)) {
    if (g->kind != SPT_GEOM_BEM_WELL && g->kind != SPT_GEOM_BEM_POST)
      return fail(c, SPT_EINVAL, "the image-charge BEM needs a BEM geometry");
    if (g->image == SPT_IMAGE_BEM_MIRROR && g->kind != SPT_GEOM_BEM_POST)
      return fail(c, SPT_EINVAL, "the mirror-doubled image surface is built for the tip");
    if (!(g->z0 > 0.0)) return fail(c, SPT_EINVAL, "the image-charge BEM needs z0 > 0");
    if (!c->img_P) return fail(c, SPT_ENOBEM, "image-charge BEM without an uploaded image table");
  }
  const bool bem = g->kind == SPT_GEOM_BEM_POST || g->kind == SPT_GEOM_BEM_WELL;
  if (bem) {
    if (!c->bem_nodes) return fail(c, SPT_ENOBEM, "BEM geometry without an uploaded table");
    if (!(g->z_handoff > 0.0)) return fail(c, SPT_EINVAL, "z_handoff must be positive");
  }
  if (g->kind == SPT_GEOM_BEM_WELL) {
    if (!(g->H > 0.0)) return fail(c, SPT_EINVAL, "well depth H must be positive");
    if (!(g->bottom_fillet >= 0.0 && g->rim_fillet >= 0.0 && g->bottom_fillet + g->rim_fillet < g->H))
      return fail(c, SPT_EINVAL, "well fillet radii must be >= 0 with sum < H");
    if (g->image != SPT_IMAGE_NONE && g->image != SPT_IMAGE_BEM)
      return fail(c, SPT_EINVAL, "the well takes no image charges or the image-charge BEM");
  }
  if (g->kind != SPT_GEOM_FLAT && !(g->R > 0.0)) return fail(c, SPT_EINVAL, "R must be positive");
  return SPT_OK;
}

int spt_geometry_field(spt_ctx* c, const spt_geometry* g, const double* pos, int64_t n, double* E,
                       int32_t precision) {
  if (!c) return SPT_EINVAL;
  int rc = check_geom(c, g);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!pos || !E))) return fail(c, SPT_EINVAL, "spt_geometry_field: bad arguments");
  if (n == 0) return SPT_OK;
  HIPCHK(c, spt_begin(c));
  const int PB = (g->kind == SPT_GEOM_BEM_POST || g->kind == SPT_GEOM_BEM_WELL) ? field_points_per_block(c, n)
                                                                                : FIELD_NMAX;
  const unsigned nb = (unsigned)((n + PB - 1) / PB);
  if (precision == SPT_PREC_MIXED)
    geometry_field_kernel<FIELD_T, FIELD_NMAX, SPT_PREC_MIXED><<<nb, FIELD_T, 0, c->stream>>>(*g, bem_view(c), pos, n,
                                                                                              E, PB);
  else
    geometry_field_kernel<FIELD_T, FIELD_NMAX, SPT_PREC_FP64><<<nb, FIELD_T, 0, c->stream>>>(*g, bem_view(c), pos, n,
                                                                                             E, PB);
  HIPCHK(c, hipGetLastError());
  return SPT_OK;
}

int spt_pair_forces(spt_ctx* c, const spt_geometry* g, const double* pos, const double* charge, const uint8_t* active,
                    const int64_t* offsets, int32_t n_groups, double a, double* Fc, double* Fi, int32_t precision) {
  if (!c) return SPT_EINVAL;
  int rc = check_geom(c, g);
  if (rc) return rc;
  if (n_groups < 0 || (n_groups > 0 && (!pos || !charge || !active || !offsets)))
    return fail(c, SPT_EINVAL, "spt_pair_forces: bad arguments");
  if (n_groups == 0) return SPT_OK;
  HIPCHK(c, spt_begin(c));
  // group sizes are checked against the block size on the host side of
  // the caller's offsets: read them back (small) to validate.
  int64_t* h = new int64_t[n_groups + 1];
  hipError_t e = hipMemcpy(h, offsets, sizeof(int64_t) * (n_groups + 1), hipMemcpyDeviceToHost);
  int64_t mx = 0;
  bool bad = false;
  for (int i = 0; e == hipSuccess && i < n_groups; ++i) {
    const int64_t s = h[i + 1] - h[i];
    if (s < 0) bad = true;
    mx = s > mx ? s : mx;
  }
  delete[] h;
  if (e != hipSuccess) return fail(c, SPT_EHIP, hipGetErrorString(e));
  if (bad || mx > 1024) return fail(c, SPT_EINVAL, "group sizes must be in [0, 1024]");
  if (precision == SPT_PREC_MIXED)
    pair_forces_kernel<1024, SPT_PREC_MIXED><<<n_groups, 1024, 0, c->stream>>>(*g, pos, charge, active, offsets, a * a,
                                                                              Fc, Fi);
  else
    pair_forces_kernel<1024, SPT_PREC_FP64><<<n_groups, 1024, 0, c->stream>>>(*g, pos, charge, active, offsets, a * a,
                                                                             Fc, Fi);
  HIPCHK(c, hipGetLastError());
  return SPT_OK;
}

int spt_kill_mask(spt_ctx* c, const spt_geometry* g, const double* pos, int64_t n, uint8_t* out) {
  if (!c) return SPT_EINVAL;
  int rc = check_geom(c, g);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!pos || !out))) return fail(c, SPT_EINVAL, "spt_kill_mask: bad arguments");
  if (n == 0) return SPT_OK;
  HIPCHK(c, spt_begin(c));
  kill_kernel<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(*g, pos, n, out);
  HIPCHK(c, hipGetLastError());
  return SPT_OK;
}

}  // extern "C"

template <int T, int NMAX, int PR, int NM = 0>
static int launch_trace(spt_ctx* c, const TraceParams& P, int n_groups) {
  int per_cu = 0;
  HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trace_kernel<T, NMAX, PR, NM>, T, 0));
  if (per_cu < 1) per_cu = 1;
  int64_t grid = (int64_t)per_cu * c->n_cu;
  if (grid > n_groups) grid = n_groups;
  trace_kernel<T, NMAX, PR, NM><<<(unsigned)grid, T, 0, c->stream>>>(P);
  HIPCHK(c, hipGetLastError());
  return SPT_OK;
}


static ImageView image_view(const spt_ctx* c) {
  ImageView B;
  B.P = c->img_P;
  B.n_modes = c->img_modes;
  B.n_quad = c->img_quad;
  B.prof = c->d_img_prof;
  B.segs = c->d_img_segs;
  B.quad = c->d_img_quadv;
  B.Ainv = c->d_img_Ainv;
  B.rho_floor = c->img_floor;
  B.d_lo = c->img_dlo;
  B.d_hi = c->img_dhi;
  B.rhs = c->d_img_rhs;
  B.sigma = c->d_img_sigma;
  B.mirror = c->d_img_mirror;
  B.mirror_mode = 0;
  B.plane_z = 0.0;
  return B;
}

static int ensure_image_scratch(spt_ctx* c, int n_groups, int64_t n) {
  if ((size_t)n_groups > c->img_groups_cap) {
    if (c->d_img_rhs) (void)hipFree(c->d_img_rhs);
    if (c->d_img_sigma) (void)hipFree(c->d_img_sigma);
    c->d_img_rhs = c->d_img_sigma = nullptr;
    c->img_groups_cap = 0;
    const size_t per = (size_t)2 * IMG_MAX_MODES * (size_t)c->img_P;
    HIPCHK(c, hipMalloc(&c->d_img_rhs, sizeof(double) * per * n_groups));
    HIPCHK(c, hipMalloc(&c->d_img_sigma, sizeof(double) * per * n_groups));
    c->img_groups_cap = n_groups;
  }
  if ((size_t)n > c->img_part_cap) {
    if (c->d_img_mirror) (void)hipFree(c->d_img_mirror);
    if (c->d_img_F) (void)hipFree(c->d_img_F);
    c->d_img_mirror = c->d_img_F = nullptr;
    c->img_part_cap = 0;
    HIPCHK(c, hipMalloc(&c->d_img_mirror, sizeof(double) * 8 * n));
    HIPCHK(c, hipMalloc(&c->d_img_F, sizeof(double) * 3 * n));
    c->img_part_cap = n;
  }
  return SPT_OK;
}

static PosView stage_pos_view(const LockParams& L) {
  PosView X;
  X.x = L.T.kbuf + (size_t)48 * L.T.n_total;
  X.y = L.T.kbuf + (size_t)49 * L.T.n_total;
  X.z = L.T.kbuf + (size_t)50 * L.T.n_total;
  X.stride = 1;
  return X;
}

// The image-charge force for every (running) group: per group one
// workgroup forms the right-hand sides and solves sigma_m = A_m^-1 rhs_m
// (img_rhs_kernel), then another the force (img_force_kernel). The scratch
// must have been sized by ensure_image_scratch. Groups that are not
// running are skipped.
static int launch_image_force(spt_ctx* c, const spt_geometry& G, PosView X, const double* charge,
                              const uint8_t* active, const int64_t* offsets, const int32_t* gflags, int running,
                              int n_groups, double a2, double* F, hipStream_t s) {
  ImageView B = image_view(c);
  if (G.image == SPT_IMAGE_BEM_MIRROR) {
    B.mirror_mode = 1;
    B.plane_z = -G.z0;
  }
  const int nm = B.n_modes;
  if (nm <= 9)
    img_rhs_kernel<9><<<n_groups, 256, 0, s>>>(B, X, charge, active, offsets, gflags, running);
  else if (nm <= 17)
    img_rhs_kernel<17><<<n_groups, 256, 0, s>>>(B, X, charge, active, offsets, gflags, running);
  else
    img_rhs_kernel<33><<<n_groups, 256, 0, s>>>(B, X, charge, active, offsets, gflags, running);
  if (nm <= 9)
    img_force_kernel<9><<<n_groups, 256, 0, s>>>(B, G, X, charge, active, offsets, gflags, running, a2, F);
  else if (nm <= 17)
    img_force_kernel<17><<<n_groups, 256, 0, s>>>(B, G, X, charge, active, offsets, gflags, running, a2, F);
  else
    img_force_kernel<33><<<n_groups, 256, 0, s>>>(B, G, X, charge, active, offsets, gflags, running, a2, F);
  HIPCHK(c, hipGetLastError());
  return SPT_OK;
}

constexpr int kLockstepUnroll = 4;  // attempts per captured graph replay

// PR: SPT_PREC_FP64, SPT_PREC_MIXED (FP32 far-field sums and pairs) or
// SPT_PREC_PAIR32 (FP64 field, FP32 pairs).
template <int TG, int PR>
static int run_lockstep(spt_ctx* c, LockParams L, int n_groups, int64_t n, FieldShape fs, int64_t max_size) {
  constexpr int FPR = PR == SPT_PREC_MIXED ? SPT_PREC_MIXED : SPT_PREC_FP64;  // field
  constexpr int APR = PR == SPT_PREC_FP64 ? SPT_PREC_FP64 : SPT_PREC_MIXED;   // pairs
  hipStream_t s = c->own;
  const int PB = fs.PB;
  const unsigned nb_field = (unsigned)((n + PB - 1) / PB);
  ls_pgroup_kernel<<<(n_groups + 255) / 256, 256, 0, s>>>(L.T.offsets, n_groups, (int32_t*)L.pgroup);
  reset_u32_kernel<<<1, 1, 0, s>>>(L.done);
  reset_u64_kernel<<<1, 1, 0, s>>>(L.T.ev_count);
  if (L.T.n_out > 0) reset_u64_kernel<<<1, 1, 0, s>>>(L.T.snap_count);
  ls_init_kernel<TG><<<n_groups, TG, 0, s>>>(L);
  const bool bin = L.X.key != nullptr;
  // groups of at most 16 share the acceleration and end-of-attempt blocks,
  // kPackThreads / 16 of them per block (16 lanes each)
  const bool pack4 = TG == 64 && max_size <= 16;
  const unsigned nb_bin = (unsigned)((n + kBinThreads - 1) / kBinThreads);
  const unsigned nb_slots = (unsigned)((n + 63) / 64);
  // blocks for the records evaluated one per thread (multipole zone, table-only tiles)
  const bool one_per_thread = L.T.bem.mp || L.T.bem.tl_tree;
  if (bin) {
    fill_u32_kernel<<<(L.X.n_keys + 255) / 256, 256, 0, s>>>(L.X.hist, L.X.n_keys, 0u);
    fill_u32_kernel<<<1, 64, 0, s>>>(L.X.n_start, 1, 0u);

    ls_bin_key_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(L);
  }
  HIPCHK(c, hipGetLastError());
  // kLockstepUnroll attempts of every running group, captured once in a HIP
  // graph and replayed.
  auto enqueue = [&]() -> int {
    for (int u = 0; u < kLockstepUnroll; ++u) {
      if (bin) {
        ls_bin_count_kernel<<<nb_bin, kBinThreads, 0, s>>>(L);
        ls_bin_scan_kernel<<<1, kBinThreads, 0, s>>>(L);
        ls_bin_slot_kernel<<<nb_bin, kBinThreads, 0, s>>>(L);
      }
      for (int st = 0; st < 7; ++st) {
        if (bin) {
          const unsigned nb_one = one_per_thread ? (unsigned)((n + fs.T - 1) / fs.T) : 0u;
          if (fs.T == 64)
            ls_bin_field_kernel<64><<<nb_slots + nb_one, 64, 0, s>>>(L, (int)nb_slots, st);
          else if (fs.T == 128)
            ls_bin_field_kernel<128><<<nb_slots + nb_one, 128, 0, s>>>(L, (int)nb_slots, st);
          else if (fs.T == 192)
            ls_bin_field_kernel<192><<<nb_slots + nb_one, 192, 0, s>>>(L, (int)nb_slots, st);
          else
            ls_bin_field_kernel<256><<<nb_slots + nb_one, 256, 0, s>>>(L, (int)nb_slots, st);
        } else if (L.sur) {
          if (FPR == SPT_PREC_FP64)
            ls_surrogate_kernel<true><<<(unsigned)((n + MLP64_PTS - 1) / MLP64_PTS), 64, 0, s>>>(L, st);
          else
            ls_surrogate_kernel<false><<<(unsigned)((n + MLP_WAVE_PTS - 1) / MLP_WAVE_PTS), 64, 0, s>>>(L, st);
        }
        if (bin) {
        } else if (fs.T == 192) {
          ls_field_kernel<192, 64, FPR><<<nb_field, 192, 0, s>>>(L, PB, st);
        } else {
          ls_field_kernel<256, 64, FPR><<<nb_field, 256, 0, s>>>(L, PB, st);
        }
        if (L.imgF) {
          const int r = launch_image_force(c, L.T.G, stage_pos_view(L), L.T.charge, L.active, L.T.offsets,
                                           L.g_flags, LS_RUNNING, n_groups, L.T.a2, (double*)L.imgF, s);
          if (r) return r;
        }
        // packed groups: stage 7's kernel also ends the attempt
        const unsigned nb_pack = (unsigned)((n_groups + kPackThreads / 16 - 1) / (kPackThreads / 16));
        constexpr int GP = kPackThreads / 16;
        if (pack4 && st == 6)
          ls_accel_kernel<kPackThreads, APR, GP, true, 6><<<nb_pack, kPackThreads, 0, s>>>(L, st);
        else if (pack4 && st == 0)
          ls_accel_kernel<kPackThreads, APR, GP, false, 0><<<nb_pack, kPackThreads, 0, s>>>(L, st);
        else if (pack4 && st == 1)
          ls_accel_kernel<kPackThreads, APR, GP, false, 1><<<nb_pack, kPackThreads, 0, s>>>(L, st);
        else if (pack4 && st == 2)
          ls_accel_kernel<kPackThreads, APR, GP, false, 2><<<nb_pack, kPackThreads, 0, s>>>(L, st);
        else if (pack4 && st == 3)
          ls_accel_kernel<kPackThreads, APR, GP, false, 3><<<nb_pack, kPackThreads, 0, s>>>(L, st);
        else if (pack4 && st == 4)
          ls_accel_kernel<kPackThreads, APR, GP, false, 4><<<nb_pack, kPackThreads, 0, s>>>(L, st);
        else if (pack4)
          ls_accel_kernel<kPackThreads, APR, GP, false, 5><<<nb_pack, kPackThreads, 0, s>>>(L, st);
        else
          ls_accel_kernel<TG, APR><<<n_groups, TG, 0, s>>>(L, st);
      }
      if (!pack4) ls_end_kernel<TG><<<n_groups, TG, 0, s>>>(L);
    }
    return SPT_OK;
  };
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rq = enqueue();
  HIPCHK(c, hipStreamEndCapture(s, &graph));
  if (rq) {
    (void)hipGraphDestroy(graph);
    return rq;
  }
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return fail(c, SPT_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  int64_t iters = 0;
  int rc = SPT_OK;
  while (true) {
    hipError_t e = hipGraphLaunch(exec, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->h_done, L.done, sizeof(unsigned int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      rc = fail(c, SPT_EHIP, std::string("lockstep replay: ") + hipGetErrorString(e));
      break;
    }
    iters += kLockstepUnroll;
    if (*c->h_done >= (unsigned)n_groups) break;
    if (iters > L.T.max_steps + 2 * kLockstepUnroll) {
      rc = fail(c, SPT_EHIP, "lockstep schedule did not terminate");
      break;
    }
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  return rc;
}

// Buffers of the binned gather for n particles (BinView).
static int ensure_bin_scratch(spt_ctx* c, int64_t n) {
  if (!c->d_bin_ctl) {
    HIPCHK(c, hipMalloc(&c->d_bin_ctl, sizeof(unsigned) * (2 * kBinMaxKeys + 64)));
  }
  if ((size_t)n > c->bin_cap) {
    if (c->d_bin_key) (void)hipFree(c->d_bin_key);
    if (c->d_bin_slot) (void)hipFree(c->d_bin_slot);
    if (c->d_bin_src) (void)hipFree(c->d_bin_src);
    if (c->d_bin_rec) (void)hipFree(c->d_bin_rec);
    if (c->d_bin_glist) (void)hipFree(c->d_bin_glist);
    c->d_bin_key = c->d_bin_slot = c->d_bin_glist = nullptr;
    c->d_bin_src = nullptr;
    c->d_bin_rec = nullptr;
    c->bin_cap = 0;
    HIPCHK(c, hipMalloc(&c->d_bin_key, sizeof(int32_t) * n));
    HIPCHK(c, hipMalloc(&c->d_bin_slot, sizeof(int32_t) * n));
    HIPCHK(c, hipMalloc(&c->d_bin_glist, sizeof(int32_t) * n));
    HIPCHK(c, hipMalloc(&c->d_bin_src, n));
    HIPCHK(c, hipMalloc(&c->d_bin_rec, (2 * sizeof(double2) + 2 * sizeof(int32_t)) * n));
    c->bin_cap = n;
  }
  return SPT_OK;
}

static int ensure_lockstep_scratch(spt_ctx* c, int64_t n, int n_groups) {
  if (!c->own) {
    HIPCHK(c, hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming));
    HIPCHK(c, hipMalloc(&c->d_done, 64));
    HIPCHK(c, hipHostMalloc(&c->h_done, 64, hipHostMallocDefault));
  }
  if ((size_t)n > c->part_cap) {
    if (c->d_pgroup) (void)hipFree(c->d_pgroup);
    if (c->d_active) (void)hipFree(c->d_active);
    if (c->d_sur) (void)hipFree(c->d_sur);
    if (c->d_fvalid) (void)hipFree(c->d_fvalid);
    c->d_pgroup = nullptr;
    c->d_active = nullptr;
    c->d_sur = nullptr;
    c->d_fvalid = nullptr;
    c->part_cap = 0;
    HIPCHK(c, hipMalloc(&c->d_pgroup, sizeof(int32_t) * n));
    HIPCHK(c, hipMalloc(&c->d_active, n));
    HIPCHK(c, hipMalloc(&c->d_sur, n));
    HIPCHK(c, hipMalloc(&c->d_fvalid, n));
    c->part_cap = n;
  }
  if ((size_t)n_groups > c->group_cap) {
    if (c->d_gdbl) (void)hipFree(c->d_gdbl);
    if (c->d_gflags) (void)hipFree(c->d_gflags);
    c->d_gdbl = nullptr;
    c->d_gflags = nullptr;
    c->group_cap = 0;
    HIPCHK(c, hipMalloc(&c->d_gdbl, sizeof(double) * 5 * n_groups));
    HIPCHK(c, hipMalloc(&c->d_gflags, sizeof(int32_t) * n_groups));
    c->group_cap = n_groups;
  }
  return SPT_OK;
}

extern "C" {

static int trace_impl(spt_ctx* c, const spt_trace_args* a, bool no_alloc) {
  if (!c || !a) return SPT_EINVAL;
  int rc = check_geom(c, &a->geom);
  if (rc) return rc;
  if (a->n_groups < 0 || a->n_particles < 0) return fail(c, SPT_EINVAL, "negative sizes");
  if (a->n_screens < 0 || a->n_screens > SPT_MAX_SCREENS) return fail(c, SPT_EINVAL, "too many screens");
  if (a->n_groups == 0) return SPT_OK;
  if (!a->group_offsets || !a->pos || !a->vel || !a->t_birth || !a->charge || !a->mass || !a->t_death ||
      !a->attempts || !a->accepted || !a->active_stages || !a->margin || !a->status || !a->event_count ||
      (a->event_capacity > 0 && !a->events))
    return fail(c, SPT_EINVAL, "spt_trace: null pointer");
  if (!(a->rtol >= 0.0) || !(a->atol >= 0.0) || !(a->max_steps > 0)) return fail(c, SPT_EINVAL, "bad tolerances");
  if (a->n_out < 0 || (a->n_out > 0 && (!a->t_out || !a->snapshot_count || (a->snapshot_capacity > 0 && !a->snapshots))))
    return fail(c, SPT_EINVAL, "spt_trace: output times need t_out, snapshots and snapshot_count");
  HIPCHK(c, spt_begin(c));
  // validate the CSR offsets and find the largest group
  int64_t* h = new int64_t[a->n_groups + 1];
  hipError_t e = hipMemcpy(h, a->group_offsets, sizeof(int64_t) * (a->n_groups + 1), hipMemcpyDeviceToHost);
  int64_t mx = 0;
  bool bad = e != hipSuccess || h[0] != 0 || h[a->n_groups] != a->n_particles;
  for (int i = 0; !bad && i < a->n_groups; ++i) {
    const int64_t s = h[i + 1] - h[i];
    if (s < 1) bad = true;
    mx = s > mx ? s : mx;
  }
  delete[] h;
  if (e != hipSuccess) return fail(c, SPT_EHIP, hipGetErrorString(e));
  if (bad) return fail(c, SPT_EINVAL, "group_offsets must start at 0, end at n_particles, with sizes >= 1");
  int T = a->block_threads ? a->block_threads : 256;
  if (T != 64 && T != 128 && T != 256 && T != 512 && T != 1024)
    return fail(c, SPT_EINVAL, "block_threads must be 64, 128, 256, 512 or 1024");
  const int nmax_of_T = T <= 256 ? T : (T == 512 ? 256 : 64);
  if (mx > nmax_of_T) {
    if (a->block_threads) return fail(c, SPT_EINVAL, "largest group exceeds what block_threads supports");
    T = mx <= 64 ? 64 : (mx <= 128 ? 128 : 256);
    if (mx > 256) return fail(c, SPT_EINVAL, "groups larger than 256 particles are not supported");
  }
  if (a->mode != SPT_MODE_PERSISTENT && a->mode != SPT_MODE_LOCKSTEP) return fail(c, SPT_EINVAL, "unknown mode");
  if (a->precision != SPT_PREC_FP64 && a->precision != SPT_PREC_MIXED && a->precision != SPT_PREC_PAIR32)
    return fail(c, SPT_EINVAL, "unknown precision");
  if (a->mode == SPT_MODE_PERSISTENT) {
    if (a->precision == SPT_PREC_PAIR32 && T > 256)
      return fail(c, SPT_EINVAL, "precision PAIR32 in persistent mode is built for 64-, 128- and 256-thread workgroups");
    if (is_image_bem(a->geom.image)) {
      if (a->precision == SPT_PREC_MIXED)
        return fail(c, SPT_EINVAL, "the image-charge BEM in persistent mode runs with precision FP64 or PAIR32");
      if (a->block_threads && a->block_threads != 256)
        return fail(c, SPT_EINVAL, "the image-charge BEM in persistent mode runs in 256-thread workgroups");
      T = 256;
    }
  }
  if (a->use_surrogate) {
    if (a->mode != SPT_MODE_LOCKSTEP) return fail(c, SPT_EINVAL, "the field surrogate runs in lockstep mode only");
    if (!c->have_sur) return fail(c, SPT_EINVAL, "use_surrogate set but no surrogate (spt_surrogate_set)");
    if (a->geom.kind != SPT_GEOM_BEM_POST)
      return fail(c, SPT_EINVAL, "the field surrogate is built for the post-tip geometry");
  }
  if (a->precision == SPT_PREC_MIXED && a->mode == SPT_MODE_PERSISTENT && T != 256)
    return fail(c, SPT_EINVAL, "mixed precision in persistent mode is built for 256-thread workgroups only");
  if (a->field_order && a->mode == SPT_MODE_LOCKSTEP) {
    // every index the field gather takes must be a particle, each once
    std::vector<int32_t> ord((size_t)a->n_particles);
    HIPCHK(c, hipMemcpy(ord.data(), a->field_order, sizeof(int32_t) * ord.size(), hipMemcpyDeviceToHost));
    std::vector<unsigned char> seen(ord.size(), 0);
    for (int32_t k : ord) {
      if (k < 0 || (int64_t)k >= a->n_particles || seen[k])
        return fail(c, SPT_EINVAL, "spt_trace: field_order must be a permutation of 0 .. n_particles - 1");
      seen[k] = 1;
    }
  }
  const size_t need = (size_t)(a->mode == SPT_MODE_LOCKSTEP ? kSlotsLockstep : kSlots) * (size_t)a->n_particles;
  if (no_alloc) {
    if (a->mode != SPT_MODE_PERSISTENT) return fail(c, SPT_EINVAL, "spt_trace_noalloc runs the persistent mode only");
    if (!c->static_scratch) return fail(c, SPT_EHIP, "spt_trace_noalloc: no device scratch (" + c->static_err + ")");
    if (need > kStaticKbufElems)
      return fail(c, SPT_EINVAL, "spt_trace_noalloc: n_particles " + std::to_string(a->n_particles) +
                                     " exceeds the static scratch's " +
                                     std::to_string(kStaticKbufElems / kSlots) + " (64 MiB)");
  } else if (need > c->kbuf_elems) {
    if (c->d_kbuf) (void)hipFree(c->d_kbuf);
    c->d_kbuf = nullptr;
    c->kbuf_elems = 0;
    HIPCHK(c, hipMalloc(&c->d_kbuf, sizeof(double) * need));
    c->kbuf_elems = need;
  }
  TraceParams P;
  memset(&P, 0, sizeof(P));
  P.G = a->geom;
  P.bem = bem_view(c);
  P.SP.n = a->n_screens;
  for (int i = 0; i < a->n_screens; ++i) P.SP.z[i] = a->screens[i];
  P.n_groups = a->n_groups;
  P.n_total = a->n_particles;
  P.offsets = a->group_offsets;
  P.order = a->group_order;
  P.pos = a->pos;
  P.vel = a->vel;
  P.t_birth = a->t_birth;
  P.charge = a->charge;
  P.mass = a->mass;
  P.t_death = a->t_death;
  P.kbuf = no_alloc ? c->static_scratch + kStaticKbufOffset : c->d_kbuf;
  P.q_over_m = a->q_over_m;
  P.a2 = a->plummer_radius * a->plummer_radius;
  P.t_max = a->t_max;
  P.rtol = a->rtol;
  P.atol = a->atol;
  P.safety = a->safety;
  P.fmin_ = a->min_factor;
  P.fmax_ = a->max_factor;
  P.init_frac = a->initial_step_fraction;
  P.max_steps = a->max_steps;
  P.attempts = a->attempts;
  P.accepted = a->accepted;
  P.active_stages = a->active_stages;
  P.margin = a->margin;
  P.status = a->status;
  P.events = a->events;
  P.ev_cap = a->event_capacity;
  P.ev_count = (unsigned long long*)a->event_count;
  P.queue = no_alloc ? reinterpret_cast<int*>(c->static_scratch) : c->d_queue;
  P.t_out = a->t_out;
  P.n_out = a->n_out;
  P.snaps = a->snapshots;
  P.snap_cap = a->snapshot_capacity;
  P.snap_count = (unsigned long long*)a->snapshot_count;
  P.t_end = a->t_end;
  if (a->mode == SPT_MODE_LOCKSTEP) {
    int rc2 = ensure_lockstep_scratch(c, a->n_particles, a->n_groups);
    if (rc2) return rc2;
    LockParams L;
    L.T = P;
    L.pgroup = c->d_pgroup;
    L.active = c->d_active;
    L.g_t = c->d_gdbl;
    L.g_h = c->d_gdbl + a->n_groups;
    L.g_hr = c->d_gdbl + 2 * (size_t)a->n_groups;
    L.g_next = c->d_gdbl + 3 * (size_t)a->n_groups;
    L.g_nact = c->d_gdbl + 4 * (size_t)a->n_groups;
    L.g_flags = c->d_gflags;
    L.done = c->d_done;
    L.sur = a->use_surrogate ? c->d_sur : nullptr;
    L.fvalid = c->d_fvalid;
    L.order = a->field_order;
    memset(&L.X, 0, sizeof(L.X));
    const bool bin = c->bin_mode != 0 && (a->geom.kind == SPT_GEOM_BEM_POST || a->geom.kind == SPT_GEOM_BEM_WELL) &&
                     P.bem.coarse && !a->use_surrogate && a->precision != SPT_PREC_MIXED && P.bem.n_seg >= 1;
    if (bin) {
      int rb = ensure_bin_scratch(c, a->n_particles);
      if (rb) return rb;
      const int n_tile = P.bem.tl_tree ? c->tl_n : 0;  // kTileMaxTiles at most
      int shift = 0;
      while (kBinBands * ((P.bem.n_seg + (1 << shift) - 1) >> shift) + n_tile + kBinMpKeys > kBinMaxKeys) ++shift;
      L.X.nsk = (P.bem.n_seg + (1 << shift) - 1) >> shift;
      L.X.shift = shift;
      L.X.key_tile0 = kBinBands * L.X.nsk;
      L.X.key_mp0 = L.X.key_tile0 + n_tile;
      L.X.n_keys = L.X.key_mp0 + (P.bem.mp ? kBinMpKeys : 0);
      L.X.key = c->d_bin_key;
      L.X.slot = c->d_bin_slot;
      L.X.src = c->d_bin_src;
      L.X.hist = c->d_bin_ctl;
      L.X.cursor = c->d_bin_ctl + kBinMaxKeys;
      L.X.total = c->d_bin_ctl + 2 * kBinMaxKeys;
      L.X.gcount = c->d_bin_ctl + 2 * kBinMaxKeys + 8;  // 7 stage counts
      L.X.list_on = c->d_bin_ctl + 2 * kBinMaxKeys + 16;
      L.X.n_start = c->d_bin_ctl + 2 * kBinMaxKeys + 17;
      L.X.glist = c->d_bin_glist;
      L.X.key_tab0 = L.X.key_tile0 + (P.bem.tl_tree ? c->tl_nwin : 0);
      L.X.rin = c->d_bin_rec;
      L.X.rout = c->d_bin_rec + a->n_particles;
      L.X.rtag = reinterpret_cast<int32_t*>(c->d_bin_rec + 2 * a->n_particles);
      L.X.rkey = L.X.rtag + a->n_particles;
      L.X.append_only = c->bin_mode == 2;
    }
    memset(&L.V, 0, sizeof(L.V));
    memset(&L.V64, 0, sizeof(L.V64));
    L.imgF = nullptr;
    if (is_image_bem(a->geom.image)) {
      int rc4 = ensure_image_scratch(c, a->n_groups, a->n_particles);
      if (rc4) return rc4;
      L.imgF = c->d_img_F;
    }
    if (a->use_surrogate) {
      L.V.cfg = c->sur_cfg;
      L.V.p = c->d_sur_p;
      L.V.split = c->d_sur_split;
      L.V64.cfg = c->sur_cfg;
      L.V64.p = c->d_sur_p;
      L.V64.frag = c->d_sur_frag;
    }
    const FieldShape fs = field_shape(c, a->n_particles,
                                      a->geom.kind == SPT_GEOM_BEM_POST || a->geom.kind == SPT_GEOM_BEM_WELL);
    HIPCHK(c, hipEventRecord(c->ev_in, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->own, c->ev_in, 0));
    int rc3;
    const int pr = a->precision;
#define SPT_LS(TG)                                                                                         \
  (pr == SPT_PREC_MIXED    ? run_lockstep<TG, SPT_PREC_MIXED>(c, L, a->n_groups, a->n_particles, fs, mx)  \
   : pr == SPT_PREC_PAIR32 ? run_lockstep<TG, SPT_PREC_PAIR32>(c, L, a->n_groups, a->n_particles, fs, mx) \
                           : run_lockstep<TG, SPT_PREC_FP64>(c, L, a->n_groups, a->n_particles, fs, mx))
    if (mx <= 64)
      rc3 = SPT_LS(64);
    else if (mx <= 128)
      rc3 = SPT_LS(128);
    else
      rc3 = SPT_LS(256);
#undef SPT_LS
    HIPCHK(c, hipEventRecord(c->ev_out, c->own));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_out, 0));
    return rc3;
  }
  const bool p32 = a->precision == SPT_PREC_PAIR32;
  if (is_image_bem(a->geom.image)) {
    if (no_alloc) return fail(c, SPT_EINVAL, "spt_trace_noalloc does not run the image-charge BEM");
    int ri = ensure_image_scratch(c, a->n_groups, a->n_particles);
    if (ri) return ri;
    P.IB = image_view(c);
    if (a->geom.image == SPT_IMAGE_BEM_MIRROR) {
      P.IB.mirror_mode = 1;
      P.IB.plane_z = -a->geom.z0;
    }
  }
  reset_int_kernel<<<1, 1, 0, c->stream>>>(P.queue, 0);
  reset_u64_kernel<<<1, 1, 0, c->stream>>>(P.ev_count);
  if (P.n_out > 0) reset_u64_kernel<<<1, 1, 0, c->stream>>>(P.snap_count);
  if (is_image_bem(a->geom.image)) {
    if (c->img_modes <= 17)
      return p32 ? launch_trace<256, 256, SPT_PREC_PAIR32, 17>(c, P, a->n_groups)
                 : launch_trace<256, 256, SPT_PREC_FP64, 17>(c, P, a->n_groups);
    return p32 ? launch_trace<256, 256, SPT_PREC_PAIR32, 33>(c, P, a->n_groups)
               : launch_trace<256, 256, SPT_PREC_FP64, 33>(c, P, a->n_groups);
  }
  switch (T) {
    case 64:
      return p32 ? launch_trace<64, 64, SPT_PREC_PAIR32>(c, P, a->n_groups)
                 : launch_trace<64, 64, SPT_PREC_FP64>(c, P, a->n_groups);
    case 128:
      return p32 ? launch_trace<128, 128, SPT_PREC_PAIR32>(c, P, a->n_groups)
                 : launch_trace<128, 128, SPT_PREC_FP64>(c, P, a->n_groups);
    case 256:
      return a->precision == SPT_PREC_MIXED ? launch_trace<256, 256, SPT_PREC_MIXED>(c, P, a->n_groups)
             : p32                          ? launch_trace<256, 256, SPT_PREC_PAIR32>(c, P, a->n_groups)
                                            : launch_trace<256, 256, SPT_PREC_FP64>(c, P, a->n_groups);
    case 512: return launch_trace<512, 256, SPT_PREC_FP64>(c, P, a->n_groups);
    default: return launch_trace<1024, 64, SPT_PREC_FP64>(c, P, a->n_groups);
  }
}

int spt_trace(spt_ctx* c, const spt_trace_args* a) { return trace_impl(c, a, false); }

int spt_trace_noalloc(spt_ctx* c, const spt_trace_args* a) { return trace_impl(c, a, true); }

}  // extern "C"

// ============================================================ surrogate host
template <int TA, int TB, int EPI, bool SPLIT>
static void mlp_gemm(hipStream_t s, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
                     int ldc, const float* bias, const float* Zin, float* Zout, int ldz, float* kpart) {
  const int tiles = (M / 16) * (N / 16);
  int ks = 1;
  if (kpart)
    while (ks < 64 && tiles * ks < 2048 && (K / 32) % (2 * ks) == 0 && K / (2 * ks) >= 64) ks *= 2;
  const dim3 grid((unsigned)((tiles + 3) / 4), (unsigned)ks);
  if (ks == 1) {
    mlp_gemm_kernel<TA, TB, EPI, SPLIT><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, bias, Zin, Zout, ldz);
  } else {  // split-K partials (STORE epilogue, C contiguous with ldc == N), then the ordered sum
    mlp_gemm_kernel<TA, TB, EPI, SPLIT><<<grid, 256, 0, s>>>(M, N, K / ks, A, lda, B, ldb, kpart, N, bias, Zin,
                                                             Zout, ldz);
    mlp_ksum_kernel<<<(unsigned)(((int64_t)M * N + 255) / 256), 256, 0, s>>>(kpart, ks, (int64_t)M * N, C);
  }
}

constexpr int kMlpKsMax = 64;

struct MlpTrainBufs {
  float* X0;
  float* Tb;
  float* Z[MLP_DEPTH];
  float* H[MLP_DEPTH];
  float* Y;
  float* dY;
  float* dZa;
  float* dZb;
  float* part;
  float* grad;
  float* kpart;
  float* loss;
};

static size_t mlp_train_floats(int B) {
  return (size_t)B * MLP_IN + (size_t)B * 4 + 2 * MLP_DEPTH * (size_t)B * MLP_W + 2 * (size_t)B * MLP_OUTP +
         2 * (size_t)B * MLP_W + (size_t)(B + 255) / 256 + 64 + MLP_NPARAM + 64 +
         (size_t)kMlpKsMax * MLP_W * MLP_W + 64;
}

static MlpTrainBufs mlp_train_bufs(float* base, int B) {
  MlpTrainBufs T;
  float* q = base;
  auto take = [&](size_t n) {
    float* r = q;
    q += (n + 63) / 64 * 64;  // 256-byte aligned pieces
    return r;
  };
  T.X0 = take((size_t)B * MLP_IN);
  T.Tb = take((size_t)B * 3);
  for (int l = 0; l < MLP_DEPTH; ++l) T.Z[l] = take((size_t)B * MLP_W);
  for (int l = 0; l < MLP_DEPTH; ++l) T.H[l] = take((size_t)B * MLP_W);
  T.Y = take((size_t)B * MLP_OUTP);
  T.dY = take((size_t)B * MLP_OUTP);
  T.dZa = take((size_t)B * MLP_W);
  T.dZb = take((size_t)B * MLP_W);
  T.part = take((size_t)(B + 255) / 256);
  T.grad = take(MLP_NPARAM);
  T.loss = take(1);
  T.kpart = take((size_t)kMlpKsMax * MLP_W * MLP_W);
  return T;
}

static int ensure_mlp_train(spt_ctx* c, int B) {
  const size_t need = mlp_train_floats(B) + 64 * 32;
  if (need > c->train_floats) {
    if (c->d_train) HIPCHK(c, hipFree(c->d_train));
    c->d_train = nullptr;
    c->train_floats = 0;
    HIPCHK(c, hipMalloc(&c->d_train, sizeof(float) * need));
    c->train_floats = need;
  }
  return SPT_OK;
}

// Forward, loss and backward of one batch (X0, Tb already built); the
// gradient lands in T.grad and the loss at loss_dst (device).
template <bool SPLIT>
static void mlp_step(hipStream_t s, const MlpTrainBufs& T, const float* p, int B, const float* w, float* loss_dst) {
  mlp_gemm<0, 1, EPI_BIAS_SILU, SPLIT>(s, B, MLP_W, MLP_IN, T.X0, MLP_IN, p + MLP_OFF_W1, MLP_IN, T.H[0], MLP_W,
                                       p + MLP_OFF_B1, nullptr, T.Z[0], MLP_W, nullptr);
  for (int l = 1; l < MLP_DEPTH; ++l) {
    const int64_t o = MLP_OFF_H + (l - 1) * MLP_HSTRIDE;
    mlp_gemm<0, 1, EPI_BIAS_SILU, SPLIT>(s, B, MLP_W, MLP_W, T.H[l - 1], MLP_W, p + o, MLP_W, T.H[l], MLP_W,
                                         p + o + (int64_t)MLP_W * MLP_W, nullptr, T.Z[l], MLP_W, nullptr);
  }
  mlp_gemm<0, 1, EPI_BIAS, SPLIT>(s, B, MLP_OUTP, MLP_W, T.H[MLP_DEPTH - 1], MLP_W, p + MLP_OFF_WO, MLP_W, T.Y,
                                  MLP_OUTP, p + MLP_OFF_BO, nullptr, nullptr, 0, nullptr);
  const int nb = (B + 255) / 256;
  mlp_loss_kernel<<<nb, 256, 0, s>>>(T.Y, T.Tb, B, w[0], w[1], w[2], T.dY, T.part);
  mlp_loss_final_kernel<<<1, 64, 0, s>>>(T.part, nb, B, loss_dst);
  float* g = T.grad;
  mlp_gemm<1, 0, EPI_STORE, SPLIT>(s, MLP_OUTP, MLP_W, B, T.dY, MLP_OUTP, T.H[MLP_DEPTH - 1], MLP_W, g + MLP_OFF_WO,
                                   MLP_W, nullptr, nullptr, nullptr, 0, T.kpart);
  mlp_colsum_kernel<<<MLP_OUTP, 256, 0, s>>>(T.dY, B, MLP_OUTP, g + MLP_OFF_BO);
  mlp_gemm<0, 0, EPI_DSILU, SPLIT>(s, B, MLP_W, MLP_OUTP, T.dY, MLP_OUTP, p + MLP_OFF_WO, MLP_W, T.dZa, MLP_W, nullptr,
                                   T.Z[MLP_DEPTH - 1], nullptr, MLP_W, nullptr);
  float* da = T.dZa;
  float* db = T.dZb;
  for (int l = MLP_DEPTH - 1; l >= 1; --l) {
    const int64_t o = MLP_OFF_H + (l - 1) * MLP_HSTRIDE;
    mlp_gemm<1, 0, EPI_STORE, SPLIT>(s, MLP_W, MLP_W, B, da, MLP_W, T.H[l - 1], MLP_W, g + o, MLP_W, nullptr, nullptr,
                                     nullptr, 0, T.kpart);
    mlp_colsum_kernel<<<MLP_W, 256, 0, s>>>(da, B, MLP_W, g + o + (int64_t)MLP_W * MLP_W);
    mlp_gemm<0, 0, EPI_DSILU, SPLIT>(s, B, MLP_W, MLP_W, da, MLP_W, p + o, MLP_W, db, MLP_W, nullptr, T.Z[l - 1],
                                     nullptr, MLP_W, nullptr);
    float* t = da;
    da = db;
    db = t;
  }
  mlp_gemm<1, 0, EPI_STORE, SPLIT>(s, MLP_W, MLP_IN, B, da, MLP_W, T.X0, MLP_IN, g + MLP_OFF_W1, MLP_IN, nullptr,
                                   nullptr, nullptr, 0, T.kpart);
  mlp_colsum_kernel<<<MLP_W, 256, 0, s>>>(da, B, MLP_W, g + MLP_OFF_B1);
}

static int check_mlp_cfg(spt_ctx* c, const spt_mlp_cfg* cfg) {
  if (!cfg) return fail(c, SPT_EINVAL, "null surrogate config");
  if (!(cfg->r_unit > 0.0) || !(cfg->env_r0 > 0.0) || !(cfg->env_pow >= 0.0))
    return fail(c, SPT_EINVAL, "surrogate config needs r_unit > 0, env_r0 > 0, env_pow >= 0");
  for (int i = 0; i < 3; ++i)
    if (!(cfg->out_scale[i] != 0.0)) return fail(c, SPT_EINVAL, "surrogate out_scale must be nonzero");
  return SPT_OK;
}

extern "C" {

int spt_bem_potential(spt_ctx* c, const double* pos, int64_t n, double* phi) {
  if (!c) return SPT_EINVAL;
  if (!c->bem_nodes) return fail(c, SPT_ENOBEM, "no BEM table uploaded");
  if (n < 0 || (n > 0 && (!pos || !phi))) return fail(c, SPT_EINVAL, "spt_bem_potential: bad arguments");
  if (n == 0) return SPT_OK;
  HIPCHK(c, spt_begin(c));
  potential_kernel<<<(unsigned)((n + 127) / 128), 128, 0, c->stream>>>(c->d_prof, c->d_sigma, c->bem_nodes, 1.0, 20,
                                                                       pos, n, phi);
  HIPCHK(c, hipGetLastError());
  return SPT_OK;
}

int spt_mlp_nparams(int64_t* n) {
  if (!n) return SPT_EINVAL;
  *n = MLP_NPARAM;
  return SPT_OK;
}

int spt_mlp_grad(spt_ctx* c, const spt_mlp_cfg* cfg, const float* params, const double* x, const double* y,
                 int32_t batch, const float* loss_weights, int32_t split, float* grad, float* loss) {
  if (!c) return SPT_EINVAL;
  int rc = check_mlp_cfg(c, cfg);
  if (rc) return rc;
  if (!params || !x || !y || !loss_weights || !grad || !loss) return fail(c, SPT_EINVAL, "spt_mlp_grad: null pointer");
  if (batch < 32 || batch % 32) return fail(c, SPT_EINVAL, "batch must be a positive multiple of 32");
  HIPCHK(c, spt_begin(c));
  rc = ensure_mlp_train(c, batch);
  if (rc) return rc;
  const MlpTrainBufs T = mlp_train_bufs(c->d_train, batch);
  hipStream_t s = c->stream;
  mlp_batch_kernel<<<(batch + 255) / 256, 256, 0, s>>>(*cfg, x, y, batch, batch, 0ULL, 0, T.X0, T.Tb);
  if (split)
    mlp_step<true>(s, T, params, batch, loss_weights, T.loss);
  else
    mlp_step<false>(s, T, params, batch, loss_weights, T.loss);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(grad, T.grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToDevice, s));
  HIPCHK(c, hipMemcpyAsync(loss, T.loss, sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  return SPT_OK;
}

int spt_mlp_train(spt_ctx* c, const spt_mlp_cfg* cfg, float* params, float* adam_m, float* adam_v, int64_t adam_t0,
                  const double* x, const double* y, int64_t n_samples, int32_t batch, int32_t steps, float lr,
                  uint64_t seed, const float* loss_weights, int32_t split, float* loss_out) {
  if (!c) return SPT_EINVAL;
  int rc = check_mlp_cfg(c, cfg);
  if (rc) return rc;
  if (!params || !adam_m || !adam_v || !x || !y || !loss_weights || (steps > 0 && !loss_out))
    return fail(c, SPT_EINVAL, "spt_mlp_train: null pointer");
  if (batch < 32 || batch % 32) return fail(c, SPT_EINVAL, "batch must be a positive multiple of 32");
  if (n_samples < 1 || steps < 0 || adam_t0 < 0) return fail(c, SPT_EINVAL, "spt_mlp_train: bad sizes");
  if (steps == 0) return SPT_OK;
  HIPCHK(c, spt_begin(c));
  rc = ensure_mlp_train(c, batch);
  if (rc) return rc;
  const MlpTrainBufs T = mlp_train_bufs(c->d_train, batch);
  hipStream_t s = c->stream;
  float* d_loss = nullptr;
  HIPCHK(c, hipMalloc(&d_loss, sizeof(float) * steps));
  HIPCHK(c, hipGetLastError());
  const double b1 = 0.9, b2 = 0.999;
  for (int i = 0; i < steps; ++i) {
    const int64_t t = adam_t0 + i + 1;
    mlp_batch_kernel<<<(batch + 255) / 256, 256, 0, s>>>(*cfg, x, y, n_samples, batch, seed, t, T.X0, T.Tb);
    if (split)
      mlp_step<true>(s, T, params, batch, loss_weights, d_loss + i);
    else
      mlp_step<false>(s, T, params, batch, loss_weights, d_loss + i);
    const float c1 = (float)(1.0 - pow(b1, (double)t)), c2 = (float)(1.0 - pow(b2, (double)t));
    mlp_adam_kernel<<<(unsigned)((MLP_NPARAM + 255) / 256), 256, 0, s>>>(params, T.grad, adam_m, adam_v, MLP_NPARAM,
                                                                         lr, (float)b1, (float)b2, 1e-8f, c1, c2);
    if (i == 0) {
      const hipError_t e0 = hipGetLastError();
      if (e0 != hipSuccess) {
        (void)hipFree(d_loss);
        return fail(c, SPT_EHIP, std::string("spt_mlp_train: first step launch: ") + hipGetErrorString(e0));
      }
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    (void)hipFree(d_loss);
    return fail(c, SPT_EHIP, std::string("spt_mlp_train: step launches: ") + hipGetErrorString(e));
  }
  if (e == hipSuccess) e = hipMemcpyAsync(loss_out, d_loss, sizeof(float) * steps, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_loss);
  if (e != hipSuccess) return fail(c, SPT_EHIP, std::string("spt_mlp_train: ") + hipGetErrorString(e));
  return SPT_OK;
}

int spt_mlp_infer(spt_ctx* c, const spt_mlp_cfg* cfg, const float* params, const double* pos, int64_t n,
                  int32_t precision, double* out) {
  if (!c) return SPT_EINVAL;
  int rc = check_mlp_cfg(c, cfg);
  if (rc) return rc;
  if (!params || n < 0 || (n > 0 && (!pos || !out))) return fail(c, SPT_EINVAL, "spt_mlp_infer: bad arguments");
  if (precision != SPT_PREC_FP64 && precision != SPT_PREC_MIXED) return fail(c, SPT_EINVAL, "unknown precision");
  if (n == 0) return SPT_OK;
  HIPCHK(c, spt_begin(c));
  hipStream_t s = c->stream;
  if (precision == SPT_PREC_FP64) {
    if (!c->d_mlp_frag) HIPCHK(c, hipMalloc(&c->d_mlp_frag, sizeof(double) * MLP64_FRAG_N));
    mlp_frag64_kernel<<<(unsigned)((MLP64_FRAG_N + 255) / 256), 256, 0, s>>>(params, c->d_mlp_frag);
    Mlp64View V;
    V.cfg = *cfg;
    V.p = params;
    V.frag = c->d_mlp_frag;
    mlp_infer64_kernel<<<(unsigned)((n + MLP64_PTS - 1) / MLP64_PTS), 64, 0, s>>>(V, pos, n, out);
  } else {
    if (!c->d_mlp_split) HIPCHK(c, hipMalloc(&c->d_mlp_split, sizeof(__bf16) * 3 * MLP_NPARAM));
    mlp_split_kernel<<<(unsigned)((MLP_NPARAM + 255) / 256), 256, 0, s>>>(params, MLP_NPARAM, c->d_mlp_split);
    MlpView V;
    V.cfg = *cfg;
    V.p = params;
    V.split = c->d_mlp_split;
    mlp_infer_kernel<<<(unsigned)((n + MLP_WAVE_PTS - 1) / MLP_WAVE_PTS), 64, 0, s>>>(V, pos, n, out);
  }
  HIPCHK(c, hipGetLastError());
  return SPT_OK;
}

int spt_surrogate_set(spt_ctx* c, const spt_mlp_cfg* cfg, const float* params) {
  if (!c) return SPT_EINVAL;
  if (!params) {
    c->have_sur = false;
    return SPT_OK;
  }
  int rc = check_mlp_cfg(c, cfg);
  if (rc) return rc;
  if (!(cfg->band >= 0.0)) return fail(c, SPT_EINVAL, "surrogate band must be >= 0");
  HIPCHK(c, spt_begin(c));
  if (!c->d_sur_p) HIPCHK(c, hipMalloc(&c->d_sur_p, sizeof(float) * MLP_NPARAM));
  if (!c->d_sur_split) HIPCHK(c, hipMalloc(&c->d_sur_split, sizeof(__bf16) * 3 * MLP_NPARAM));
  if (!c->d_sur_frag) HIPCHK(c, hipMalloc(&c->d_sur_frag, sizeof(double) * MLP64_FRAG_N));
  HIPCHK(c, hipMemcpyAsync(c->d_sur_p, params, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToDevice, c->stream));
  mlp_split_kernel<<<(unsigned)((MLP_NPARAM + 255) / 256), 256, 0, c->stream>>>(c->d_sur_p, MLP_NPARAM,
                                                                                c->d_sur_split);
  mlp_frag64_kernel<<<(unsigned)((MLP64_FRAG_N + 255) / 256), 256, 0, c->stream>>>(c->d_sur_p, c->d_sur_frag);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->sur_cfg = *cfg;
  c->have_sur = true;
  return SPT_OK;
}

}  // extern "C"

extern "C" {

int spt_toroidal_q(spt_ctx* c, const double* chi, int64_t n, int32_t n_max, double* q, double* dq) {
  if (!c) return SPT_EINVAL;
  if (n < 0 || n_max < 0 || n_max > SPT_TOROIDAL_MAX_M || (n > 0 && (!chi || !q || !dq)))
    return fail(c, SPT_EINVAL, "spt_toroidal_q: need 0 <= n_max <= 64 and device arrays");
  if (n == 0) return SPT_OK;
  HIPCHK(c, spt_begin(c));
  toroidal_q_kernel<<<(unsigned)((n + 127) / 128), 128, 0, c->stream>>>(chi, n, n_max, q, dq);
  HIPCHK(c, hipGetLastError());
  return SPT_OK;
}

int spt_ring_modes(spt_ctx* c, const double* rho, const double* z, const double* a, const double* za, int64_t n,
                   int32_t n_max, double* phi, double* dphi_drho, double* dphi_dz, double* g) {
  if (!c) return SPT_EINVAL;
  if (n < 0 || n_max < 0 || n_max > SPT_TOROIDAL_MAX_M || (n > 0 && (!rho || !z || !a || !za)))
    return fail(c, SPT_EINVAL, "spt_ring_modes: need 0 <= n_max <= 64 and device arrays");
  if (n == 0) return SPT_OK;
  HIPCHK(c, spt_begin(c));
  ring_modes_kernel<<<(unsigned)((n + 127) / 128), 128, 0, c->stream>>>(rho, z, a, za, n, n_max, phi, dphi_drho,
                                                                       dphi_dz, g);
  HIPCHK(c, hipGetLastError());
  return SPT_OK;
}

}  // extern "C"

extern "C" {

int spt_image_upload(spt_ctx* c, const double* prof, int32_t n, int32_t n_max, int32_t n_sub, double d_lo,
                     double d_hi, double* A_out) {
  if (!c) return SPT_EINVAL;
  if (!prof || n < 3 || n_max < 0 || n_max >= IMG_MAX_MODES || n_sub < 1 || n_sub > 64)
    return fail(c, SPT_EINVAL, "spt_image_upload: need n_nodes >= 3, 0 <= n_max <= 32, 1 <= n_subdiv <= 64");
  HIPCHK(c, spt_begin(c));
  for (double** p : {&c->d_img_prof, &c->d_img_segs, &c->d_img_quadv, &c->d_img_Ainv, &c->d_img_rhs,
                     &c->d_img_sigma}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  c->img_P = 0;
  c->img_groups_cap = 0;
  const int M = n_max + 1, ns = n - 1, nq = ns * n_sub * 4;
  double rmax = 0.0;
  for (int i = 0; i < n; ++i) rmax = prof[2 * i] > rmax ? prof[2 * i] : rmax;
  const double floor_ = 1e-6 * (rmax > 1e-300 ? rmax : 1e-300);
  HIPCHK(c, hipMalloc(&c->d_img_prof, sizeof(double) * 2 * n));
  HIPCHK(c, hipMalloc(&c->d_img_segs, sizeof(double) * IMG_SEG * ns));
  HIPCHK(c, hipMalloc(&c->d_img_quadv, sizeof(double) * 5 * nq));
  HIPCHK(c, hipMalloc(&c->d_img_Ainv, sizeof(double) * M * n * n));
  HIPCHK(c, hipMemcpyAsync(c->d_img_prof, prof, sizeof(double) * 2 * n, hipMemcpyHostToDevice, c->stream));
  img_seg_table_kernel<<<(ns + 255) / 256, 256, 0, c->stream>>>(c->d_img_prof, ns, c->d_img_segs);
  img_quad_kernel<<<(nq + 255) / 256, 256, 0, c->stream>>>(c->d_img_prof, ns, n_sub, c->d_img_quadv);
  double *lo = nullptr, *hi = nullptr;
  HIPCHK(c, hipMalloc(&lo, sizeof(double) * M * n * ns));
  HIPCHK(c, hipMalloc(&hi, sizeof(double) * M * n * ns));
  const int64_t pairs = (int64_t)n * ns;
  img_assemble_kernel<<<(unsigned)((pairs + 63) / 64), 64, 0, c->stream>>>(c->d_img_prof, n, M, floor_, lo, hi);
  img_combine_kernel<<<(unsigned)(((int64_t)M * n * n + 255) / 256), 256, 0, c->stream>>>(lo, hi, n, M,
                                                                                         c->d_img_Ainv);
  HIPCHK(c, hipGetLastError());
  if (A_out) {  // operators before inversion, (m, i, j) row-major
    std::vector<double> h((size_t)M * n * n);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_img_Ainv, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int m = 0; m < M; ++m)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) A_out[((size_t)m * n + i) * n + j] = h[((size_t)m * n + j) * n + i];
  }
  (void)hipFree(lo);
  (void)hipFree(hi);
  if (!c->blas) {
    if (rocblas_create_handle(&c->blas) != rocblas_status_success) return fail(c, SPT_EHIP, "rocblas_create_handle");
    rocblas_set_stream(c->blas, c->stream);
  }
  rocblas_int *ipiv = nullptr, *info = nullptr;
  HIPCHK(c, hipMalloc(&ipiv, sizeof(rocblas_int) * n));
  HIPCHK(c, hipMalloc(&info, sizeof(rocblas_int)));
  int bad = 0;
  for (int m = 0; m < M && !bad; ++m) {
    double* A = c->d_img_Ainv + (size_t)m * n * n;
    rocblas_status s1 = rocsolver_dgetrf(c->blas, n, n, A, n, ipiv, info);
    rocblas_int h_info = 0;
    HIPCHK(c, hipMemcpyAsync(&h_info, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    rocblas_status s2 = rocsolver_dgetri(c->blas, n, A, n, ipiv, info);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (s1 != rocblas_status_success || s2 != rocblas_status_success || h_info != 0) bad = 1;
  }
  (void)hipFree(ipiv);
  (void)hipFree(info);
  if (bad) return fail(c, SPT_EHIP, "image operator inversion failed (singular mode operator?)");
  c->img_P = n;
  c->img_modes = M;
  c->img_quad = nq;
  c->img_floor = floor_;
  c->img_dlo = d_lo;
  c->img_dhi = d_hi;
  return SPT_OK;
}

int spt_image_force(spt_ctx* c, const spt_geometry* g, const double* pos, const double* charge,
                    const uint8_t* active, const int64_t* offsets, int32_t n_groups, double plummer_radius,
                    double* F) {
  if (!c) return SPT_EINVAL;
  int rc = check_geom(c, g);
  if (rc) return rc;
  if (!is_image_bem(g->image))
    return fail(c, SPT_EINVAL, "spt_image_force needs image = SPT_IMAGE_BEM or SPT_IMAGE_BEM_MIRROR");
  if (n_groups < 0 || (n_groups > 0 && (!pos || !charge || !active || !offsets || !F)))
    return fail(c, SPT_EINVAL, "spt_image_force: bad arguments");
  if (n_groups == 0) return SPT_OK;
  HIPCHK(c, spt_begin(c));
  int64_t last = 0;
  std::vector<int64_t> h(n_groups + 1);
  HIPCHK(c, hipMemcpy(h.data(), offsets, sizeof(int64_t) * (n_groups + 1), hipMemcpyDeviceToHost));
  for (int i = 0; i < n_groups; ++i)
    if (h[i + 1] - h[i] < 1 || h[i + 1] - h[i] > 256)
      return fail(c, SPT_EINVAL, "groups must hold 1 to 256 particles");
  last = h[n_groups];
  rc = ensure_image_scratch(c, n_groups, last);
  if (rc) return rc;
  PosView X;
  X.x = pos;
  X.y = pos + 1;
  X.z = pos + 2;
  X.stride = 3;
  return launch_image_force(c, *g, X, charge, active, offsets, nullptr, 0, n_groups, plummer_radius * plummer_radius,
                            F, c->stream);
}

}  // extern "C"
