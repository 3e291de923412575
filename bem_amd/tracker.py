"""SpecificParticleTracer on the MI355X (reference tracker.py:42-417).

Same constructor, defaults and return types as the reference. ``run()``
hands the whole integration to one persistent HIP kernel (spt_trace):
every group advances with its own t and h, accept/reject, kill and screen
detection happen on the device, and the host only uploads the initial
state and reads back final state, counters and screen events.

Additions: ``group_sizes`` (uneven groups, a CSR layout instead of
``n_emit``-sized chunks), ``block_threads`` (workgroup size), and
``last_run`` (per-group attempt/accept counters, particle-stage counts,
device time), and ``mode``: "persistent" (one workgroup integrates whole
groups from a device queue), "lockstep" (every running group makes one
attempt per iteration, per-stage kernels replayed from a HIP graph; the
field gather spreads each particle over many threads) or "auto" (lockstep
when there are fewer groups than twice the compute units, or with FP32 pairs
or the image-charge BEM); and
``precision``: "fp64" (the plain method) or "mixed" (FP32 ring-kernel sums
over segments at least two panel lengths away and FP32 pair forces, both
Kahan-compensated; FP64 elsewhere) or "pairs32" (FP64 field, FP32 Kahan
pair forces); and ``surrogate``: a trained
bem_amd.surrogate.FieldSurrogate for the geometry, used in lockstep mode
for the BEM perturbation wherever its validated band allows; and
``far_rule``: None (the reference's fixed far rule) or (k4, k2), the
distance-adaptive far rule of spt_bem_far_rule for the BEM gathers of the
trace; and ``multipole``: None or a validated bem_amd.bem.multipole.Multipole
of the geometry, whose exterior expansion replaces the segment sums beyond
its radius; ``tiles``: None or a bem_amd.bem.tiles.Tiles of the geometry,
whose tables replace the segments outside each tile's window (field-error
budgets, not the plain method).
"""
import warnings

import numpy as np

from . import _lib
from .backend import resolve_backend
from .constants import ELEMENTARY_CHARGE, ev_c_to_si_momentum, ev_to_kg, si_momentum_to_ev_c
from .emission import group_indices, group_offsets
from .particles import ParticleGroup

DEFAULT_PLUMMER_RADIUS = 1e-12  # m, reference tracker.py:22
DEFAULT_RTOL = 1e-7             # reference tracker.py:23
DEFAULT_ATOL = 1e-12            # reference tracker.py:39
N_STAGES = 7                    # six Dormand-Prince stages + endpoint evaluation per attempt


class SpecificParticleTracer:
    """Electrons emitted in groups from a cathode, tracked with in-group
    Coulomb and image forces by a per-group adaptive RK45 (reference
    tracker.py:42-153 for the full parameter description)."""

    def __init__(self, initial_particles, n_emit, geometry, screens, t_out=None,
                 plummer_radius=DEFAULT_PLUMMER_RADIUS, backend="mi355x", n_workers=1, rtol=None, atol=None,
                 t_max=None, z_max=None, group_sizes=None, block_threads=0, device=0, max_steps=10_000_000,
                 mode="auto", precision="fp64", surrogate=None, far_rule=None, multipole=None, near_width=None,
                 clusters=None, spatial_order=False, binned=True, tiles=None):
        if n_workers != 1:
            raise ValueError(f"n_workers is only supported for backend='cpu' (got backend={backend!r}, "
                             f"n_workers={n_workers!r})")
        self.initial_particles = initial_particles
        self.n_emit = n_emit
        self.output_times_host = None if t_out is None else np.sort(np.asarray(t_out, dtype=float))
        self.screen_positions = list(screens)
        if len(self.screen_positions) > _lib.SPT_MAX_SCREENS:
            raise ValueError(f"at most {_lib.SPT_MAX_SCREENS} screens are supported")
        self.plummer_radius = plummer_radius
        self.z_max = z_max
        if z_max is not None and self.screen_positions and z_max < max(self.screen_positions):
            warnings.warn(
                f"z_max={z_max!r} is below the furthest screen ({max(self.screen_positions)!r}): particles are "
                "killed before reaching it, so crossings recorded at screens beyond z_max are spurious and their "
                "number will vary with step size. Raise z_max above every screen, or drop those screens.",
                stacklevel=2,
            )
        self.n_workers = 1
        self.backend_name = backend
        self.device = device
        if backend not in ("mi355x",):
            resolve_backend(backend)  # raises ValueError for unknown names
        from .geometry import lower

        self.geometry = geometry = lower(geometry)  # a reference geometry is rebuilt from its worker_args()
        self.rtol = rtol if rtol is not None else DEFAULT_RTOL
        self.atol = atol if atol is not None else DEFAULT_ATOL
        self.block_threads = int(block_threads)
        if mode not in ("auto", "persistent", "lockstep"):
            raise ValueError(f"mode must be 'auto', 'persistent' or 'lockstep', got {mode!r}")
        self.mode = mode
        from .geometry import _precision_code

        self.precision = precision
        self._precision = _precision_code(precision, trace=True)
        if surrogate is not None:
            if mode == "persistent":
                raise ValueError("a field surrogate runs in lockstep mode only (mode='lockstep' or 'auto')")
            if surrogate.geometry is not geometry:
                raise ValueError("the surrogate was trained for a different geometry object")
        self.surrogate = surrogate
        if far_rule is not None:
            far_rule = (float(far_rule[0]), float(far_rule[1]))
            if not (far_rule[0] >= 1.0 and far_rule[1] >= far_rule[0]):
                raise ValueError(f"far_rule needs 1 <= k4 <= k2, got {far_rule!r}")
            if geometry.bem_table() is None:
                raise ValueError("far_rule applies to BEM geometries only")
        self.far_rule = far_rule
        if (near_width is not None or clusters is not None or tiles is not None) and far_rule is None:
            raise ValueError("near_width, clusters and tiles refine the far rule: set far_rule too")
        if tiles is not None and tiles.geometry is not geometry:
            raise ValueError("the tiles were built for a different geometry object")
        self.tiles = tiles
        if clusters is not None and clusters.geometry is not geometry:
            raise ValueError("the cluster expansions were built for a different geometry object")
        self.near_width = near_width
        self.clusters = clusters
        if multipole is not None and multipole.geometry is not geometry:
            raise ValueError("the multipole expansion was fitted for a different geometry object")
        self.multipole = multipole
        self.field_shape = None  # (threads, points per block) of the field gather; None = automatic
        self.spatial_order = bool(spatial_order)
        # far-rule lockstep traces gather the field bin by bin (spt_set_lockstep_binning:
        # True/1, False/0, or 2 = every gather through the append path)
        self.binned = int(binned)
        if self.binned not in (0, 1, 2):
            raise ValueError(f"binned must be True, False or 2, got {binned!r}")
        self._field_order = None
        self.max_steps = int(max_steps)
        self._setup_state(group_sizes)
        self.t_max = t_max if t_max is not None else self._default_t_max()
        self.last_run = None

    def _setup_state(self, group_sizes):
        """Reference tracker.py:214-288, on the host, in CSR group order."""
        pg = self.initial_particles
        ids = np.asarray(pg.id)
        if group_sizes is None:
            idx = group_indices(ids, self.n_emit)
            self.n_groups, self.n_per_group = idx.shape
            order = idx.reshape(-1)
            self.offsets = np.arange(0, order.size + 1, self.n_per_group, dtype=np.int64)
        else:
            order, self.offsets = group_offsets(ids, group_sizes)
            self.n_groups = len(self.offsets) - 1
            self.n_per_group = None
        self.order = order
        mass_ev = pg.mass
        self.mass = ev_to_kg(mass_ev)
        self.charge_to_mass = float(pg.species_charge) / self.mass
        species_charge_magnitude = abs(pg.species_charge)
        weight = np.asarray(pg.weight, dtype=float)
        if not np.allclose(weight, species_charge_magnitude, rtol=1e-9, atol=0):
            warnings.warn(
                "SpecificParticleTracer models individual real particles, not statistical macroparticles: "
                "overriding initial_particles.weight (which was not uniformly "
                f"{species_charge_magnitude:.6e} C, the species' elementary charge) so every particle carries "
                "exactly one elementary charge of weight.",
                stacklevel=3,
            )
            weight = np.full_like(weight, species_charge_magnitude)
        charge = np.sign(pg.species_charge) * weight
        particle_mass = self.mass * (weight / species_charge_magnitude)
        o = order
        self.pos0 = np.ascontiguousarray(np.stack([np.asarray(pg.x)[o], np.asarray(pg.y)[o], np.asarray(pg.z)[o]], -1),
                                         dtype=np.float64)
        self.vel0 = np.ascontiguousarray(np.stack([
            ev_c_to_si_momentum(np.asarray(pg.px, dtype=float))[o] / self.mass,
            ev_c_to_si_momentum(np.asarray(pg.py, dtype=float))[o] / self.mass,
            ev_c_to_si_momentum(np.asarray(pg.pz, dtype=float))[o] / self.mass], -1))
        self.t_birth = np.ascontiguousarray(np.asarray(pg.t, dtype=float)[o])
        self.charge = np.ascontiguousarray(charge[o])
        self.particle_mass = np.ascontiguousarray(particle_mass[o])
        self.weight = np.ascontiguousarray(weight[o])
        self.ids = np.asarray(ids)[o]
        self.species = getattr(pg, "species", "electron")

    def _default_t_max(self):
        """Reference tracker.py:290-307."""
        a0 = abs(self.geometry.E_gun) * abs(self.charge_to_mass)
        zs = np.array(self.screen_positions, dtype=float)
        zp = zs[zs > 0]
        t_transit = np.sqrt(2 * zp.max() / a0) if len(zp) > 0 and a0 > 0 else 1e-12
        t_max = float(self.t_birth.max()) + 3.0 * t_transit
        if self.output_times_host is not None and len(self.output_times_host):
            t_max = max(t_max, float(self.output_times_host.max()))
        return t_max

    def _queue_order(self):
        """Largest groups first (longest-processing-time order); ties keep
        index order."""
        sizes = np.diff(self.offsets)
        return np.argsort(-sizes, kind="stable").astype(np.int32)

    def _mode_code(self, ctx):
        if self.surrogate is not None:
            return _lib.MODE_LOCKSTEP
        if self.mode == "persistent":
            return _lib.MODE_PERSISTENT
        if self.mode == "lockstep":
            return _lib.MODE_LOCKSTEP
        if self.precision == "pairs32" or getattr(self.geometry, "image_profile", None) is not None:
            return _lib.MODE_LOCKSTEP
        import ctypes

        n_cu = ctypes.c_int(0)
        ctx.check(ctx.lib.spt_device_cus(ctx.h, ctypes.byref(n_cu)), "spt_device_cus")
        return _lib.MODE_LOCKSTEP if self.n_groups < 2 * n_cu.value else _lib.MODE_PERSISTENT

    def field_order(self, ctx):
        """The lockstep field gather's particle order: particles sorted by the
        profile node closest to their initial (rho, z) (stable, so ties keep
        index order), computed once on the device. Neighbouring lanes then
        hold nearby points, so a wavefront's segment rules and cluster
        decisions agree; the field is axisymmetric, so the azimuth is
        irrelevant. None for non-BEM geometries or when switched off (the
        default: on the 262k-electron post trace it measured neutral, since
        the far-rule nodes it saves are no longer what the gather's time
        goes to)."""
        tab = self.geometry.bem_table()
        if not self.spatial_order or tab is None:
            return None
        if self._field_order is None:
            import torch

            dev = ctx.torch_device
            prof = torch.as_tensor(np.asarray(tab[0], dtype=np.float64), device=dev)
            p = torch.as_tensor(self.pos0, device=dev)
            rho, z = torch.hypot(p[:, 0], p[:, 1]), p[:, 2]
            best = torch.empty(p.shape[0], dtype=torch.int64, device=dev)
            step = max(1, (1 << 24) // prof.shape[0])
            for a in range(0, p.shape[0], step):
                d2 = (rho[a:a + step, None] - prof[None, :, 0]) ** 2 + (z[a:a + step, None] - prof[None, :, 1]) ** 2
                best[a:a + step] = torch.argmin(d2, dim=1)
            self._field_order = torch.argsort(best, stable=True).to(torch.int32).contiguous()
        return self._field_order

    def _device_inputs(self, ctx):
        """The trace's inputs in device memory, uploaded once per tracer and
        device: initial state, birth times, charges, masses, group offsets
        and queue order, output times."""
        import torch

        dev = ctx.torch_device
        key = str(dev)
        cache = getattr(self, "_dev_in", None)
        if cache is not None and cache[0] == key:
            return cache[1]

        def up(a, dt=torch.float64):
            return torch.as_tensor(a, device=dev).to(dt).contiguous()

        outs = self.output_times_host
        d = dict(pos0=up(self.pos0), vel0=up(self.vel0), tb=up(self.t_birth), q=up(self.charge),
                 m=up(self.particle_mass), offs=up(self.offsets, torch.int64),
                 qorder=up(self._queue_order(), torch.int32),
                 t_out=up(outs) if outs is not None and len(outs) else None,
                 stage_weight=up(N_STAGES * np.diff(self.offsets), torch.int64))
        self._dev_in = (key, d)
        return d

    def trace(self, event_capacity=None, no_alloc=False):
        """Run spt_trace; returns the raw device results as NumPy arrays.
        no_alloc: call spt_trace_noalloc instead (persistent mode, the
        library's static device scratch, at most 155,344 particles)."""
        return self.to_host(self.trace_device(event_capacity=event_capacity, no_alloc=no_alloc))

    @staticmethod
    def to_host(out):
        """trace_device()'s results as NumPy arrays (events and snapshots as
        structured arrays of _lib.EVENT_DTYPE)."""
        res = {k: out[k].cpu().numpy() for k in ("pos", "vel", "t_death", "attempts", "accepted", "active_stages",
                                                   "margin", "status", "t_end")}
        res.update(events=np.frombuffer(out["events"].cpu().numpy().tobytes(), dtype=_lib.EVENT_DTYPE),
                   snapshots=np.frombuffer(out["snapshots"].cpu().numpy().tobytes(), dtype=_lib.EVENT_DTYPE),
                   device_ms=out["device_ms"], particle_stages=out["particle_stages"])
        return res

    def trace_device(self, event_capacity=None, no_alloc=False):
        """spt_trace from the device-resident inputs (uploaded once per
        tracer); the results stay in device memory: torch tensors (the
        events and snapshots as raw records of _lib.EVENT_DTYPE bytes, the
        first event_count / snapshot_count of them), plus device_ms and
        particle_stages."""
        import torch

        ctx = self.geometry._ctx()
        dev = ctx.torch_device
        n = self.pos0.shape[0]
        G = self.n_groups
        d = self._device_inputs(ctx)
        pos = d["pos0"].clone()
        vel = d["vel0"].clone()
        tb, q, m, offs, qorder, t_out = d["tb"], d["q"], d["m"], d["offs"], d["qorder"], d["t_out"]
        td = torch.empty(n, dtype=torch.float64, device=dev)
        attempts = torch.zeros(G, dtype=torch.int64, device=dev)
        accepted = torch.zeros(G, dtype=torch.int64, device=dev)
        act = torch.zeros(G, dtype=torch.int64, device=dev)
        margin = torch.zeros(G, dtype=torch.float64, device=dev)
        status = torch.zeros(G, dtype=torch.int32, device=dev)
        t_end = torch.zeros(G, dtype=torch.float64, device=dev)
        ns = len(self.screen_positions)
        cap = event_capacity if event_capacity is not None else max(1024, 4 * n * max(ns, 1))
        events = torch.empty((cap, _lib.EVENT_DTYPE.itemsize), dtype=torch.uint8, device=dev)
        ev_count = torch.zeros(1, dtype=torch.int64, device=dev)
        n_out = 0 if t_out is None else int(t_out.shape[0])
        snap_cap = n * n_out  # every particle at most once per output time
        snaps = torch.empty((max(snap_cap, 1), _lib.EVENT_DTYPE.itemsize), dtype=torch.uint8, device=dev)
        snap_count = torch.zeros(1, dtype=torch.int64, device=dev)

        a = _lib.TraceArgs_t()
        a.geom = self.geometry.spt_geometry(self.z_max)
        a.n_groups = G
        a.block_threads = self.block_threads
        a.n_particles = n
        a.group_offsets = _lib.ptr(offs)
        a.group_order = _lib.ptr(qorder)
        a.pos, a.vel, a.t_birth, a.charge, a.mass, a.t_death = (_lib.ptr(x) for x in (pos, vel, tb, q, m, td))
        a.q_over_m = self.charge_to_mass
        a.plummer_radius = float(self.plummer_radius)
        a.t_max = float(self.t_max)
        a.rtol = float(self.rtol)
        a.atol = float(self.atol)
        a.safety, a.min_factor, a.max_factor, a.initial_step_fraction = 0.9, 0.2, 10.0, 1e-6
        a.max_steps = self.max_steps
        a.n_screens = ns
        a.mode = self._mode_code(ctx)
        a.precision = self._precision
        a.use_surrogate = 0
        if self.surrogate is not None and not self.surrogate.inert:
            self.surrogate.install(ctx)
            a.use_surrogate = 1
        for i, z in enumerate(self.screen_positions):
            a.screens[i] = float(z)
        a.attempts, a.accepted, a.active_stages = _lib.ptr(attempts), _lib.ptr(accepted), _lib.ptr(act)
        a.margin, a.status, a.t_end = _lib.ptr(margin), _lib.ptr(status), _lib.ptr(t_end)
        a.events = _lib.ptr(events)
        a.event_capacity = cap
        a.event_count = _lib.ptr(ev_count)
        a.t_out = _lib.ptr(t_out)
        a.n_out = n_out
        a.snapshots = _lib.ptr(snaps)
        a.snapshot_capacity = snap_cap
        a.snapshot_count = _lib.ptr(snap_count)
        a.field_order = _lib.ptr(self.field_order(ctx)) if a.mode == _lib.MODE_LOCKSTEP else None

        ctx.bind_stream()
        shape = getattr(self, "field_shape", None) or (0, 0)
        ctx.check(ctx.lib.spt_set_field_shape(ctx.h, int(shape[0]), int(shape[1])), "spt_set_field_shape")
        ctx.check(ctx.lib.spt_set_lockstep_binning(ctx.h, self.binned), "spt_set_lockstep_binning")
        k4, k2 = self.far_rule or (0.0, 0.0)
        ctx.check(ctx.lib.spt_bem_far_rule(ctx.h, k4, k2), "spt_bem_far_rule")
        if self.near_width is not None:
            ctx.check(ctx.lib.spt_bem_near_width(ctx.h, float(self.near_width)), "spt_bem_near_width")
        if self.clusters is not None:
            self.clusters.install(ctx)
        if self.tiles is not None:
            self.tiles.install(ctx)
        if self.multipole is not None and not self.multipole.inert:
            self.multipole.install(ctx)
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        try:
            if no_alloc:
                ctx.check(ctx.lib.spt_trace_noalloc(ctx.h, a), "spt_trace_noalloc")
            else:
                ctx.check(ctx.lib.spt_trace(ctx.h, a), "spt_trace")
        finally:
            ctx.lib.spt_set_field_shape(ctx.h, 0, 0)
            ctx.lib.spt_set_lockstep_binning(ctx.h, 1)
            ctx.lib.spt_bem_far_rule(ctx.h, 0.0, 0.0)
            ctx.lib.spt_bem_near_width(ctx.h, 0.0)
            ctx.lib.spt_bem_clusters(ctx.h, None, 0, None, None, None, 0, 0.0, 0.0)
            ctx.lib.spt_bem_tiles_enable(ctx.h, 0)
            ctx.lib.spt_bem_multipole(ctx.h, None, 0, 0.0, 0.0)
        t1.record()
        ctx.sync()
        torch.cuda.synchronize(dev)
        counts = torch.stack([ev_count[0], snap_count[0], torch.sum(attempts * d["stage_weight"])]).cpu()
        n_ev, n_sn, stages = (int(v) for v in counts)
        if n_ev > cap:
            return self.trace_device(event_capacity=n_ev, no_alloc=no_alloc)
        return dict(pos=pos, vel=vel, t_death=td, attempts=attempts, accepted=accepted, active_stages=act,
                    margin=margin, status=status, t_end=t_end, events=events[:n_ev],
                    snapshots=snaps[:min(n_sn, snap_cap)], device_ms=t0.elapsed_time(t1), particle_stages=stages)

    def run(self, verbose=False):
        """Reference tracker.py:312-386: (screens, trajectories) as lists of
        ParticleGroup; screen crossings in the reference's recording order
        (step batch, monotone interval, group, particle)."""
        out = self.trace()
        if np.any(out["status"] == _lib.STATUS_STEPCAP):
            raise RuntimeError(f"batched RK45 did not converge within {self.max_steps} steps")
        stalled = np.nonzero(out["status"] == _lib.STATUS_STALLED)[0]
        if len(stalled):
            raise RuntimeError(f"batched RK45 did not converge: the step size of group(s) {stalled[:8].tolist()} "
                               "collapsed below 1e-14 |t| after a rejected attempt")
        self.last_run = out
        if verbose:
            print(f"{int(out['accepted'].max())} accepted step-batches across {self.n_groups} groups "
                  f"(backend={self.backend_name})")
        ev = out["events"]
        screens = []
        for si in range(len(self.screen_positions)):
            e = ev[ev["screen"] == si]
            e = e[np.lexsort((e["particle"], e["group"], e["interval"], e["batch"]))]
            screens.append(self._build_particle_group(e))
        trajectories = []
        if self.output_times_host is not None:
            sn = out["snapshots"]
            for k in range(len(self.output_times_host)):
                e = sn[sn["screen"] == k]
                e = e[np.lexsort((e["particle"], e["group"], e["batch"]))]
                trajectories.append(self._build_particle_group(e))
        return screens, trajectories

    def _build_particle_group(self, e):
        """Reference tracker.py:388-417."""
        if len(e) == 0:
            data = dict(x=np.array([]), y=np.array([]), z=np.array([]), px=np.array([]), py=np.array([]),
                        pz=np.array([]), t=np.array([]), weight=np.array([]), status=np.array([]),
                        id=np.array([], dtype=int), species=self.species)
            return ParticleGroup(data=data)
        gidx = self.offsets[e["group"]] + e["particle"]
        v = e["vel"]
        data = dict(x=e["pos"][:, 0].copy(), y=e["pos"][:, 1].copy(), z=e["pos"][:, 2].copy(),
                    px=si_momentum_to_ev_c(v[:, 0] * self.mass), py=si_momentum_to_ev_c(v[:, 1] * self.mass),
                    pz=si_momentum_to_ev_c(v[:, 2] * self.mass), t=e["t"].copy(), weight=self.weight[gidx],
                    status=np.ones(len(e)), id=self.ids[gidx].astype(int), species=self.species)
        return ParticleGroup(data=data)


__all__ = ["SpecificParticleTracer", "DEFAULT_PLUMMER_RADIUS", "DEFAULT_RTOL", "DEFAULT_ATOL", "ELEMENTARY_CHARGE"]
