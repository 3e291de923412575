/* spt.h -- C ABI of libspt_hip.so, the MI355X (gfx950) implementation of
 * the specific_particle_tracer tracing path.
 *
 * The reference (AdamCBartnik/BEM, package specific_particle_tracer) is pure
 * Python with no FFI; its plug-in surfaces are the backend name
 * (backend.py:18-41), the Geometry ABC (geometry.py:37-94) and the
 * integrator callback protocol (batched_rk45.py:55-148). Each entry point
 * below names the reference function it replaces. A Python package binds
 * them with ctypes (bem_amd/_lib.py).
 *
 * Conventions
 *  - Pointers documented "device" are caller-owned, contiguous device
 *    buffers (e.g. torch.Tensor.data_ptr()); "host" pointers are ordinary
 *    host memory. The library owns only its context, its scratch and the
 *    BEM tables uploaded into it, and frees none of the caller's memory.
 *  - Positions are (n, 3) row-major float64 in metres, SI units throughout.
 *  - Every call returns SPT_OK (0) or a negative code; spt_last_error()
 *    gives the message. No C++ exception crosses the boundary.
 *  - One context per device. A context is not re-entrant: one host thread
 *    uses it at a time. Work is enqueued on the context's stream;
 *    spt_sync() waits for it.
 */
#ifndef SPT_H
#define SPT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPT_OK 0
#define SPT_EINVAL (-1)
#define SPT_ENOMEM (-2)
#define SPT_EHIP (-3)
#define SPT_ENOBEM (-4)
#define SPT_ESTEPCAP (-5)

#define SPT_MAX_SCREENS 16

/* Geometry kinds (reference geometry.py:97-256, bem/geometry.py:58-507). */
#define SPT_GEOM_FLAT 0     /* FlatCathode: GunField, plane image            */
#define SPT_GEOM_TIP 1      /* HemisphericalTip: closed-form sphere field    */
#define SPT_GEOM_BEM_POST 2 /* HemisphericalTipBEMGeometry (H = 0) and the   */
                            /* hemisphere-capped post (H > 0): uniform field */
                            /* + uploaded BEM perturbation, smoothstep blend */
                            /* to the flat cathode between z_handoff/2 and   */
                            /* z_handoff                                     */
#define SPT_GEOM_BEM_WELL 3 /* CylindricalWellBEMGeometry: well of radius R, */
                            /* depth H (bottom at z = -H), optional rounded  */
                            /* bottom/rim corners; uploaded BEM perturbation */
                            /* of the real truncated-plane surface, same     */
                            /* blend; killed only inside the conductor       */

/* Image-charge models (reference forces.py:60-191). */
#define SPT_IMAGE_NONE 0
#define SPT_IMAGE_PLANE 1  /* image at z' = -2 z0 - z, charge -q            */
#define SPT_IMAGE_TIP3 2   /* sphere R - z0 + plane z0: three images        */
#define SPT_IMAGE_BEM 3    /* well or tip: image-charge BEM on the uploaded */
                           /* truncated-plane image surface near, plane     */
                           /* image z0 far, blended (spt_image_upload;      */
                           /* LOCKSTEP traces only)                         */
#define SPT_IMAGE_BEM_MIRROR 4 /* tip: as SPT_IMAGE_BEM on the uploaded     */
                           /* mirror-doubled image surface, with a phantom  */
                           /* -q at z' = -2 z0 - z per particle and its     */
                           /* plane image (reference bem/geometry.py:       */
                           /* 346-427)                                      */

typedef struct spt_ctx spt_ctx;

typedef struct spt_geometry {
  int32_t kind;        /* SPT_GEOM_*                                         */
  int32_t image;       /* SPT_IMAGE_*                                        */
  double E_gun;        /* asymptotic Ez [V/m]                                */
  double R;            /* tip / post / well radius [m]                       */
  double H;            /* post height [m] (cap centre at z = H); 0 = bare    */
                       /* hemisphere; well depth [m]                         */
  double z0;           /* image-plane recession [m]                          */
  double z_handoff;    /* BEM-to-flat blend height [m] (BEM kinds)           */
  int32_t has_kill_z;  /* kill at z <= kill_z_below                          */
  int32_t has_z_max;   /* kill at z >= z_max                                 */
  double kill_z_below;
  double z_max;
  double bottom_fillet; /* well corner rounding radii [m] (0 = sharp)       */
  double rim_fillet;
} spt_geometry;

/* One screen crossing (reference screens.py:18-33,96-115). `batch` is the
 * group's attempt index of the accepted step that produced it, which is
 * the reference's loop-iteration index; sorting by (batch, interval, group,
 * particle) within a screen gives the reference's recording order. */
typedef struct spt_event {
  int32_t group;
  int32_t particle; /* index within the group */
  int32_t screen;
  int32_t interval; /* 0..2: monotone piece of the cubic Hermite */
  int64_t batch;
  double t;
  double pos[3];
  double vel[3];
} spt_event;

/* Execution modes of spt_trace.
 *  PERSISTENT: one persistent kernel; each workgroup integrates whole
 *    groups taken from a device queue (largest first). Best with many
 *    groups: no lockstep, no per-stage launches.
 *  LOCKSTEP: the reference's own schedule (every running group makes one
 *    attempt per iteration), as per-stage kernels replayed from a captured
 *    HIP graph; the field gather spreads each particle over many threads,
 *    so a few groups still fill the GPU. spt_trace blocks until done. */
#define SPT_MODE_PERSISTENT 0
#define SPT_MODE_LOCKSTEP 1

/* Arithmetic of the field gather and pair forces.
 *  FP64: everything FP64 (the plain method; parity with the reference).
 *  MIXED: ring-kernel sums over far nodes (segments at least two panel
 *    lengths away) and pair forces in FP32 with Kahan-compensated
 *    accumulation; FP64 for coordinate differences, segments within two
 *    panel lengths (including the log-singular sinh rule), the error norm,
 *    step control and the collocation solve. */
#define SPT_PREC_FP64 0
#define SPT_PREC_MIXED 1
#define SPT_PREC_PAIR32 2 /* spt_trace, LOCKSTEP: FP64 field, FP32 + Kahan pairs */

/* Fourier-feature MLP field surrogate (no counterpart in the reference; the
 * reference evaluates the BEM field directly). Fixed architecture: 32
 * frequencies -> 64 features -> 4 x (128, SiLU) -> 3 outputs (phi, E_rho,
 * E_z of the BEM perturbation), parameters FP32, flat layout [W1 (128x64),
 * b1, W2, b2, W3, b3, W4, b4, Wout (32x128, rows 3.. zero), bout (32)].
 *
 * Inputs: polar coordinates about (rho, z) = (0, center_z):
 * r = |(rho, z - center_z)|, u = (log(r / r_unit), (z - center_z) / r),
 * normalised un = (u - u_offset) * u_scale; features sin/cos(2 pi B un).
 * Outputs: physical = network output * out_scale * env(r),
 * env(r) = (1 + r / env_r0)^-env_pow. */
#define SPT_MLP_FOURIER 32
#define SPT_MLP_WIDTH 128
#define SPT_MLP_DEPTH 4

typedef struct spt_mlp_cfg {
  float fourier[2 * SPT_MLP_FOURIER]; /* frequency matrix B (32 x 2), row-major  */
  double center_z, r_unit;
  double u_offset[2], u_scale[2];
  double out_scale[3];
  double env_r0, env_pow;
  double rho_max, z_min, z_max;       /* box where spt_trace may use it,         */
  double cos_min;                     /* cone (z - center_z) / r >= cos_min,     */
  double band;                        /* and at least this far from the          */
                                      /* conductor surface (capsule of the       */
                                      /* geometry); blended to the BEM between   */
                                      /* band and 2 band                         */
} spt_mlp_cfg;

/* Group status words written by spt_trace.
 *  DONE: reached t_max.
 *  STEPCAP: made max_steps attempts first (the reference raises
 *    RuntimeError at max_steps, batched_rk45.py:320-321).
 *  STALLED: an attempt was rejected and the next step size was not above
 *    1e-14 |t| (or was zero or NaN): the group can no longer advance and
 *    stops there instead of spinning to max_steps (no counterpart in the
 *    reference; the Python layer raises the reference's RuntimeError). */
#define SPT_STATUS_DONE 0
#define SPT_STATUS_STEPCAP 1
#define SPT_STATUS_STALLED 2

typedef struct spt_trace_args {
  spt_geometry geom;
  int32_t n_groups;
  int32_t block_threads;        /* 0 = library default                      */
  int64_t n_particles;
  const int64_t* group_offsets; /* device, n_groups + 1 (CSR)               */
  const int32_t* group_order;   /* device, n_groups: queue order, or NULL   */
  double* pos;                  /* device (n, 3), in/out                    */
  double* vel;                  /* device (n, 3), in/out                    */
  const double* t_birth;        /* device (n)                               */
  const double* charge;         /* device (n) signed charge [C]             */
  const double* mass;           /* device (n) particle mass [kg]            */
  double* t_death;              /* device (n), out (inf = alive)            */
  double q_over_m;              /* species charge/mass for the field term   */
  double plummer_radius;
  double t_max;
  double rtol, atol;
  double safety, min_factor, max_factor, initial_step_fraction;
  int64_t max_steps;
  int32_t n_screens;
  int32_t mode;                 /* SPT_MODE_*                               */
  int32_t precision;            /* SPT_PREC_*                               */
  int32_t use_surrogate;        /* 1: field from the surrogate set with     */
                                /* spt_surrogate_set where it applies       */
                                /* (LOCKSTEP mode only), BEM elsewhere      */
  double screens[SPT_MAX_SCREENS];
  /* outputs, device */
  int64_t* attempts;            /* per group                                */
  int64_t* accepted;            /* per group                                */
  int64_t* active_stages;       /* per group: 7 x active particles summed   */
  double* margin;               /* per group: min |error_norm - 1|          */
  int32_t* status;              /* per group: SPT_STATUS_*                  */
  spt_event* events;            /* capacity entries                         */
  int64_t event_capacity;
  uint64_t* event_count;        /* total found (may exceed capacity)        */
  /* trajectory snapshots at output times (reference tracker.py t_out,
   * batched_rk45.py:181-209,295-312): steps land on each time; every
   * particle born by then is recorded (screen field = output index,
   * batch = attempt index, -1 at a group's start) */
  const double* t_out;          /* ascending, device; NULL when n_out = 0   */
  int32_t n_out;
  int32_t pad1;
  spt_event* snapshots;         /* device, capacity records                 */
  int64_t snapshot_capacity;
  uint64_t* snapshot_count;     /* device, total recorded                   */
  double* t_end;                /* device, per group: t after the group's   */
                                /* last attempt (t_max when done), or NULL  */
  const int32_t* field_order;   /* device, n_particles, LOCKSTEP: the order */
                                /* in which the field gather takes the      */
                                /* particles (a permutation; a spatial sort */
                                /* keeps each wavefront's points together), */
                                /* or NULL for index order. It changes a    */
                                /* fixed-rule result in the last bits only  */
                                /* (through the slice count of its block)   */
} spt_trace_args;

/* Context. */
int spt_create(int device, spt_ctx** out);
void spt_destroy(spt_ctx* ctx);
const char* spt_last_error(const spt_ctx* ctx);
int spt_set_stream(spt_ctx* ctx, void* hip_stream);
int spt_sync(spt_ctx* ctx);
int spt_device_cus(spt_ctx* ctx, int* n_cu);

/* BEM tables. Replaces AxisymmetricBEMSolution.quadrature/_prepare_field_data
 * (reference bem/axisymmetric.py:305-338,464-498): builds the fixed
 * Gauss-4 node table (rho, z, area*sigma) and per-segment geometry on the
 * device. profile_rz and sigma are host arrays of n_nodes entries. */
int spt_bem_upload(spt_ctx* ctx, const double* profile_rz, int32_t n_nodes, const double* sigma,
                   int32_t n_subdiv);

/* Distance-adaptive far rule for every later field gather on this context
 * (field evaluation, traces). No counterpart in the reference, whose far
 * rule is always the fixed sub-panel table of spt_bem_upload (reference
 * bem/axisymmetric.py:560-564): with k4 > 0, a segment at distance D >= L
 * from the point is summed over the whole segment by Gauss-8 below k4 L,
 * Gauss-4 below k2 L and Gauss-2 beyond; closer than L the near rule
 * applies as before. k4 = 0 restores the reference rule (the default). Needs
 * 1 <= k4 <= k2 (k2 = INFINITY: no Gauss-2). A field-error budget, for the
 * traces that state one, not a parity mode. */
int spt_bem_far_rule(spt_ctx* ctx, double k4, double k2);

/* Exterior multipole expansion of the uploaded BEM charge for every later
 * field gather on this context (field evaluation, traces). No counterpart
 * in the reference, which sums every segment at every point (reference
 * bem/axisymmetric.py:560-605). moments[l] = sum over the quadrature nodes
 * of (area * sigma)_q (r_q / r_ref)^l P_l(z_q / r_q), r_q = |(rho_q, z_q)|,
 * for l < n_terms (host array); r_ref >= every r_q. Points at distance
 * >= r_min from the origin take the perturbation
 *   phi = (1/4 pi) sum_l moments[l] (r_ref / r)^l P_l(cos theta) / r
 * (its field) instead of the segment sums. n_terms = 0 turns it off (the
 * default); spt_bem_upload also turns it off. Needs 1 <= n_terms <= 1024,
 * 0 < r_ref < r_min. A field-error budget, for the traces that state one
 * (r_min and n_terms come from a validation against the segment sums), not
 * a parity mode. */
int spt_bem_multipole(spt_ctx* ctx, const double* moments, int32_t n_terms, double r_ref, double r_min);

/* Axis-centred expansions of runs of consecutive segments for every later
 * far-rule gather on this context (no counterpart in the reference, which
 * sums every segment at every point, reference bem/axisymmetric.py:560-605;
 * only active while spt_bem_far_rule is). Cluster c holds segments
 * seg_begin[c] .. seg_begin[c+1] - 1 (seg_begin[0] = 0, seg_begin[n] = the
 * segment count); moments[c * n_terms + l] = sum over its fixed-rule
 * quadrature nodes of (area * sigma)_q (d_q / r_c)^l P_l((z_q - z_c) / d_q),
 * d_q = |(rho_q, z_q - z_c)| <= r_c. A point at distance r >= kappa r_c
 * from (0, z_c) takes the cluster's perturbation from
 *   phi = (1/4 pi) sum_l moments[l] (r_c / r)^l P_l(cos theta) / r
 * about (0, z_c), truncated after the first l whose tail bound
 * sum_{k > l} (k+1) q^k (q = r_c / r) is <= tol, instead of the cluster's
 * segments. Host arrays; 1 <= n_clusters <= 64, 1 <= n_terms <= 256,
 * kappa > 1, tol > 0; n_clusters = 0 turns it off (spt_bem_upload does
 * too). A field-error budget (bem_amd.bem.clusters validates it), not a
 * parity mode. */
int spt_bem_clusters(spt_ctx* ctx, const int32_t* seg_begin, int32_t n_clusters, const double* z_c,
                     const double* r_c, const double* moments, int32_t n_terms, double kappa, double tol);

/* Interval width (in the sinh-substituted variable) of the near rule in
 * far-rule gathers on this context; the reference's rule, used by every
 * fixed-rule gather, is 0.5 (reference bem/axisymmetric.py:583). 0 restores
 * 0.5; at most 4. A field-error budget, not a parity mode. */
int spt_bem_near_width(spt_ctx* ctx, double width);

/* Chebyshev tiles for every later far-rule gather on this context (no
 * counterpart in the reference, which sums every segment at every point,
 * reference bem/axisymmetric.py:541-605; only active while
 * spt_bem_far_rule is). A quadtree over the square [x0, x0 + size) x
 * [z0, z0 + size) of (rho, z): tree[k] >= 0 is the index of node k's first
 * child (four in a row: low rho low z, high rho low z, low rho high z, high
 * rho high z; always after k), -1 a cell without a tile, <= -2 tile
 * -2 - tree[k]. A point in tile t sums the segments win[2t] .. win[2t+1] - 1
 * with the far-rule gather's rules and takes the fixed-rule field of every
 * other segment from the tile's table, coef[t * 288 ..]: 12 x 12 Chebyshev
 * coefficients c[a][b] of E_rho, then of E_z, in T_a(u) T_b(v) with (u, v)
 * the point's position in the cell scaled to [-1, 1]^2 (first index a:
 * rho). The window must hold every segment within a panel length of the
 * cell (bem_amd.bem.tiles builds and validates such tables). Host arrays,
 * copied; 1 <= n_tiles <= 6144. Enables the tiles; spt_bem_upload drops
 * them. A field-error budget, not a parity mode. */
int spt_bem_tiles(spt_ctx* ctx, const int32_t* tree, int32_t n_tree, const int32_t* win, const double* coef,
                  int32_t n_tiles, double x0, double z0, double size);

/* Turns the uploaded tiles off (0) or on again (1) for later far-rule
 * gathers, keeping the upload. */
int spt_bem_tiles_enable(spt_ctx* ctx, int32_t on);

/* The field a tile tabulates, for building tables: at each of n device
 * points pts (rho, z pairs) the fixed-rule perturbation field (E_rho, E_z
 * pairs into device E) of every segment outside [win[2i], win[2i+1])
 * (device int32 pairs), summed over their sub-panel Gauss nodes as the
 * reference sums segments at least a panel length away (reference
 * bem/axisymmetric.py:560-564). */
int spt_bem_remainder(spt_ctx* ctx, const double* pts, const int32_t* win, int64_t n, double* E);

/* Collocation matrix of the P1 hat basis (reference
 * bem/axisymmetric.py:275-302,430-441), row-major into host A_out
 * (n_nodes x n_nodes). */
int spt_bem_assemble(spt_ctx* ctx, const double* profile_rz, int32_t n_nodes, double refine_ratio,
                     int32_t max_depth, double* A_out);

/* AxisymmetricBEMSolution.solve (reference bem/axisymmetric.py:411-443) for
 * Dirichlet data E_gun * z: device assembly plus an FP64 LU solve; sigma_out
 * is host (n_nodes). */
int spt_bem_solve(spt_ctx* ctx, const double* profile_rz, int32_t n_nodes, double E_gun, double refine_ratio,
                  int32_t max_depth, double* sigma_out);

/* evaluate_axisymmetric_field (reference bem/axisymmetric.py:501-605) on the
 * uploaded table: perturbation field at n device points into device E. */
int spt_bem_pert_field(spt_ctx* ctx, const double* pos, int64_t n, double* E, int32_t precision);

/* AxisymmetricBEMSolution.potential (reference bem/axisymmetric.py:445-462,
 * 275-302): perturbation potential of the uploaded BEM table at n device
 * points into device phi; per segment, adaptive bisection (split while
 * length > distance to the midpoint, depth <= 20) with Gauss-4 leaves and
 * sigma linear along the segment. Surrogate training labels. */
int spt_bem_potential(spt_ctx* ctx, const double* pos, int64_t n, double* phi);

/* Geometry.field (reference geometry.py:144-145,227-228;
 * bem/geometry.py:319-324): full external field [V/m] at device points. */
int spt_geometry_field(spt_ctx* ctx, const spt_geometry* geom, const double* pos, int64_t n, double* E,
                       int32_t precision);

/* Pair forces within CSR groups (reference forces.py:32-191; Geometry.image_force):
 * coulomb_out gets the Plummer-softened in-group Coulomb force, image_out the
 * image-charge force of geom->image. Either output may be NULL. active is
 * one byte per particle. All device. */
int spt_pair_forces(spt_ctx* ctx, const spt_geometry* geom, const double* pos, const double* charge,
                    const uint8_t* active, const int64_t* group_offsets, int32_t n_groups,
                    double plummer_radius, double* coulomb_out, double* image_out, int32_t precision);

/* Geometry.kill_mask (+ z_max, reference geometry.py:304-322): one byte per
 * device point. */
int spt_kill_mask(spt_ctx* ctx, const spt_geometry* geom, const double* pos, int64_t n, uint8_t* out);

/* ImageChargeBEMSolution.solve (reference bem/image_charge.py:469-493,
 * 538-566): upload the image surface profile (host, n_nodes x 2), assemble
 * the per-mode single-layer operators m = 0..n_max (n_max <= 32) on the
 * device and invert them (rocSOLVER getrf/getri); mirror-charge blend
 * distances d_lo < d_hi [m] and the field-quadrature n_subdiv are kept
 * for spt_image_force / spt_trace. A_out (host, optional): the operators
 * before inversion, (n_max + 1) x n x n row-major. */
int spt_image_upload(spt_ctx* ctx, const double* profile_rz, int32_t n_nodes, int32_t n_max, int32_t n_subdiv,
                     double d_lo, double d_hi, double* A_out);

/* CylindricalWellBEMGeometry / HemisphericalTipBEMGeometry.image_force with
 * z0 (reference bem/geometry.py:326-427,786-801, bem/image_charge.py:
 * 816-1056): per group (CSR
 * offsets, 1..256 members) the joint image-charge BEM force, blended with
 * the plane image at z0 over [z_handoff/2, z_handoff]; device arrays, F
 * (n x 3) [N]. geom->image must be SPT_IMAGE_BEM or SPT_IMAGE_BEM_MIRROR. */
int spt_image_force(spt_ctx* ctx, const spt_geometry* geom, const double* pos, const double* charge,
                    const uint8_t* active, const int64_t* group_offsets, int32_t n_groups, double plummer_radius,
                    double* F);

/* Launch shape of the BEM field gather (no counterpart in the reference;
 * set by the autotuner in bem_amd/autotune.py): threads per block 64, 128,
 * 192 or 256 and points per block (a power of two <= 64); 0 keeps the
 * automatic rule (about three resident blocks per compute unit, whole
 * rounds). 64 and 128 threads apply to the binned lockstep gather and the
 * FP64 spt_bem_pert_field (one or two slices per point at 64 points per
 * block); the other gathers take 256 for them. */
int spt_set_field_shape(spt_ctx* ctx, int32_t threads, int32_t points_per_block);

/* Binned gather of far-rule lockstep traces (no counterpart in the
 * reference, whose field evaluation takes the particles in index order,
 * bem/axisymmetric.py:501-605): 1 (the default) bins the particles that
 * need the gather by distance band and closest segment at the start of
 * every attempt and gathers them bin by bin from packed slot records; 0
 * gathers in particle order (field_order applies only then); 2 gives no
 * slots at attempt starts, so every gather takes the path of a particle
 * that needs one without a slot (slots in atomic order: a test that results
 * do not depend on slot order). Used when the far rule is set and the trace
 * runs neither the surrogate nor MIXED precision. Results do not depend on
 * it beyond the last bits (0 against 1 or 2); 1 and 2 agree bit for bit. */
int spt_set_lockstep_binning(spt_ctx* ctx, int32_t on);

/* toroidal_Q (reference bem/toroidal.py:188-219): Q_{m-1/2}(chi) and
 * dQ/dchi for m = 0..n_max (n_max <= 64) at n device values chi >= 1 into
 * device q, dq, each (n_max + 1) x n, mode-major. */
int spt_toroidal_q(spt_ctx* ctx, const double* chi, int64_t n, int32_t n_max, double* q, double* dq);

/* ring_field_modes and point_charge_potential_modes (reference
 * bem/ring_modes.py:92-158) for n (field, source) pairs (rho, z; a, za):
 * Phi_m, dPhi_m/drho, dPhi_m/dz of a cos(m phi') ring and the point-charge
 * coefficient g_m, m = 0..n_max, each (n_max + 1) x n mode-major; any
 * output may be NULL. */
int spt_ring_modes(spt_ctx* ctx, const double* rho, const double* z, const double* a, const double* za, int64_t n,
                   int32_t n_max, double* phi, double* dphi_drho, double* dphi_dz, double* g);

/* Surrogate training and use (no counterpart in the reference). Array
 * arguments are device pointers unless marked host.
 *  spt_mlp_nparams: number of FP32 parameters of the flat layout.
 *  spt_mlp_grad: loss and gradient (same layout as params) of the batch
 *    x (batch x 2: rho, z; FP64), y (batch x 3 physical targets: phi, E_rho,
 *    E_z; FP64); loss = mean over rows of sum_c w_c (net_c - y_c / scale_c)^2
 *    with scale_c = out_scale_c * env(r); batch a multiple of 32. Products
 *    on BF16 MFMA with FP32 accumulation: split = 0 one BF16 term per
 *    operand, split = 1 a three-term BF16 split (six products). host:
 *    loss_weights[3], loss.
 *  spt_mlp_train: `steps` Adam steps (lr, beta1 0.9, beta2 0.999, eps 1e-8,
 *    step t = adam_t0 + 1 ...) on batches drawn from the n samples by a
 *    counter-based hash of (seed, step); host: loss_weights[3], per-step
 *    losses loss_out[steps].
 *  spt_mlp_infer: (phi, E_rho, E_z) at n points (x, y, z) in physical units;
 *    precision SPT_PREC_FP64: FP64 MFMA products; SPT_PREC_MIXED: ~FP32
 *    accuracy from a three-term BF16 split (six MFMA products per tile).
 *    spt_trace uses the form matching its own precision.
 *  spt_surrogate_set: make params the surrogate that spt_trace uses when
 *    use_surrogate is set (a private copy is kept); NULL clears it. */
int spt_mlp_nparams(int64_t* n);
int spt_mlp_grad(spt_ctx* ctx, const spt_mlp_cfg* cfg, const float* params, const double* x, const double* y,
                 int32_t batch, const float* loss_weights, int32_t split, float* grad, float* loss);
int spt_mlp_train(spt_ctx* ctx, const spt_mlp_cfg* cfg, float* params, float* adam_m, float* adam_v,
                  int64_t adam_t0, const double* x, const double* y, int64_t n_samples, int32_t batch, int32_t steps,
                  float lr, uint64_t seed, const float* loss_weights, int32_t split, float* loss_out);
int spt_mlp_infer(spt_ctx* ctx, const spt_mlp_cfg* cfg, const float* params, const double* pos, int64_t n,
                  int32_t precision, double* out);
int spt_surrogate_set(spt_ctx* ctx, const spt_mlp_cfg* cfg, const float* params);

/* batched_rk45.integrate + make_accel_fn + find_crossings_grouped
 * (reference batched_rk45.py:55-401, geometry.py:265-301, screens.py:43-168,
 * tracker.py:340-372) as one persistent kernel: every group advances with
 * its own t and h, accept/reject on device, screen events appended to
 * args->events. Asynchronous; call spt_sync before reading outputs. */
int spt_trace(spt_ctx* ctx, const spt_trace_args* args);

/* spt_trace that allocates nothing. No counterpart in the reference; it is
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. PERSISTENT
 * mode only, with the same kernels and results as spt_trace in that mode;
 * the stage buffer (54 doubles per particle) and the work-queue counter
 * live in a 64 MiB file-scope device array of the library instead of
 * context memory, so n_particles <= 155,344 (larger inputs: SPT_EINVAL).
 * Calls no hipMalloc, hipFree or allocating library routine. Each device
 * holds its own copy of the array until the process ends; every call on a
 * device shares it and it keeps the last call's contents (the queue
 * counter is cleared at the start of each call), so calls that overlap on
 * one device (any context, any stream) race. */
int spt_trace_noalloc(spt_ctx* ctx, const spt_trace_args* args);

#ifdef __cplusplus
}
#endif

#endif /* SPT_H */
