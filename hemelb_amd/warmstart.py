"""Learned warm start: a 3D U-Net, trained and run by the library's HIP
kernels (``include/hemelb_unet.h``), that maps a signed-distance wall field
plus iolet-parameter channels to the periodic-state velocity and pressure
of a vessel; the LB run then starts from the equilibrium of that field.

Input channels (per voxel of a cubic crop): signed distance to the wall in
voxels (positive in the fluid) divided by 16, 10 (tau - 1/2), period / 1000
steps, half the Stokes-layer thickness, and cos / sin of the phase of the
cycle.  Output channels: the velocity in units of half the inlet density
amplitude and (density - 1) in units of that amplitude.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _native as N

CHANNELS = (16, 32, 64, 128)
IN_CHANNELS = 6
OUT_CHANNELS = 4

_SIGS = {
    "hlbn_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                              C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_void_p)]),
    "hlbn_destroy": (None, [C.c_void_p]),
    "hlbn_param_count": (C.c_int64, [C.c_void_p]),
    "hlbn_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hlbn_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hlbn_get_grads": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hlbn_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
    "hlbn_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                C.POINTER(C.c_double), C.c_int32]),
    "hlbn_adam": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float,
                            C.c_float]),
    "hlbn_device_bytes": (C.c_int64, [C.c_void_p]),
    "hlbn_last_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hlbn_stream": (C.c_void_p, [C.c_void_p]),
    "hlbn_last_error": (C.c_char_p, []),
}
_bound = False


def _lib():
    global _bound
    lib = N.lib()
    if not _bound:
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _bound = True
    return lib


def _check(rc):
    if rc != 0:
        raise N.HemeLBError(rc, (_lib().hlbn_last_error() or b"").decode())


def _ptr(a):
    if hasattr(a, "data_ptr"):        # a device tensor
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


LAYERS = ["e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "e4a", "e4b",
          "d3a", "d3b", "d2a", "d2b", "d1a", "d1b"]


def layer_shapes(chans=CHANNELS, cin_padded=16):
    c1, c2, c3, c4 = chans
    ci = [cin_padded, c1, c1, c2, c2, c3, c3, c4, c4 + c3, c3, c3 + c2, c2, c2 + c1, c1]
    co = [c1, c1, c2, c2, c3, c3, c4, c4, c3, c3, c2, c2, c1, c1]
    return list(zip(LAYERS, ci, co))


def pack_params(p: dict, chans=CHANNELS, cout=OUT_CHANNELS) -> np.ndarray:
    """Flat parameter vector (hemelb_unet.h order) from a dict of arrays
    (name.w [27,Co,Ci], name.b, head.w [Cout,C1], head.b); the first layer's
    weights are zero-padded to 16 input channels."""
    parts = []
    for name, ci, co in layer_shapes(chans):
        w = np.asarray(p[name + ".w"], np.float32)
        if w.shape[2] < ci:
            w = np.concatenate([w, np.zeros((27, co, ci - w.shape[2]), np.float32)], 2)
        parts += [w.reshape(-1), np.asarray(p[name + ".b"], np.float32)]
    parts += [np.asarray(p["head.w"], np.float32).reshape(-1),
              np.asarray(p["head.b"], np.float32)]
    return np.concatenate(parts)


def unpack(v: np.ndarray, chans=CHANNELS, cout=OUT_CHANNELS, cin=None) -> dict:
    out, k = {}, 0
    for name, ci, co in layer_shapes(chans):
        n = 27 * co * ci
        w = v[k:k + n].reshape(27, co, ci)
        out[name + ".w"] = w[:, :, :cin] if (cin is not None and name == "e1a") else w
        k += n
        out[name + ".b"] = v[k:k + co]
        k += co
    out["head.w"] = v[k:k + cout * chans[0]].reshape(cout, chans[0])
    k += cout * chans[0]
    out["head.b"] = v[k:k + cout]
    return out


def init_params(chans=CHANNELS, cin=IN_CHANNELS, cout=OUT_CHANNELS, seed=0) -> np.ndarray:
    """He-normal convolution weights, small head, zero biases."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co in layer_shapes(chans):
        fan = 27 * (cin if name == "e1a" else ci)
        w = rng.standard_normal((27, co, ci)) * math.sqrt(2.0 / fan)
        if name == "e1a":
            w[:, :, cin:] = 0.0
        p[name + ".w"] = w.astype(np.float32)
        p[name + ".b"] = np.zeros(co, np.float32)
    p["head.w"] = (rng.standard_normal((cout, chans[0])) * math.sqrt(1.0 / chans[0])
                   ).astype(np.float32)
    p["head.b"] = np.zeros(cout, np.float32)
    return pack_params(p, chans, cout)


class UNet:
    """The U-Net on one GPU for a fixed batch of cubic crops of side `side`."""

    def __init__(self, side: int, batch: int = 1, in_channels: int = IN_CHANNELS,
                 out_channels: int = OUT_CHANNELS, chans=CHANNELS, device: int = 0):
        self.side, self.batch = side, batch
        self.cin, self.cout, self.chans = in_channels, out_channels, tuple(chans)
        ch = (C.c_int32 * 4)(*chans)
        h = C.c_void_p()
        _check(_lib().hlbn_create(side, batch, in_channels, out_channels, ch, device,
                                  C.byref(h)))
        self._h = h
        self.nparam = int(_lib().hlbn_param_count(h))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib().hlbn_destroy(self._h)
            self._h = C.c_void_p()

    def set_params(self, v: np.ndarray) -> None:
        v = np.ascontiguousarray(v, np.float32)
        assert v.size == self.nparam
        _check(_lib().hlbn_set_params(self._h, _ptr(v)))

    def params(self) -> np.ndarray:
        v = np.empty(self.nparam, np.float32)
        _check(_lib().hlbn_get_params(self._h, _ptr(v)))
        return v

    def grads(self) -> np.ndarray:
        v = np.empty(self.nparam, np.float32)
        _check(_lib().hlbn_get_grads(self._h, _ptr(v)))
        return v

    def forward(self, x, out=None, device: bool = False):
        """x: [N,S,S,S,cin] binary32 (numpy, or a device tensor with
        device=True); returns the output array (numpy) or fills `out`."""
        if device:
            _check(_lib().hlbn_forward(self._h, _ptr(x), _ptr(out) if out is not None else None,
                                       1))
            return out
        x = np.ascontiguousarray(x, np.float32)
        assert x.shape == (self.batch, self.side, self.side, self.side, self.cin), x.shape
        y = np.empty((self.batch, self.side, self.side, self.side, self.cout), np.float32)
        _check(_lib().hlbn_forward(self._h, _ptr(x), _ptr(y), 0))
        return y

    def backward(self, target, mask, channel_weights=(1.0, 1.0, 1.0, 1.0), mask_count=None,
                 device: bool = False) -> float:
        w = np.ascontiguousarray(channel_weights, np.float32)
        loss = C.c_double()
        if device:
            cnt = float(mask_count)
            _check(_lib().hlbn_backward(self._h, _ptr(target), _ptr(mask), _ptr(w), cnt,
                                        C.byref(loss), 1))
            return loss.value
        t = np.ascontiguousarray(target, np.float32)
        m = np.ascontiguousarray(mask, np.uint8)
        cnt = float(m.sum()) if mask_count is None else float(mask_count)
        _check(_lib().hlbn_backward(self._h, _ptr(t), _ptr(m), _ptr(w), cnt, C.byref(loss), 0))
        return loss.value

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0) -> None:
        _check(_lib().hlbn_adam(self._h, lr, beta1, beta2, eps, weight_decay))

    def times(self):
        f, b = C.c_double(), C.c_double()
        _check(_lib().hlbn_last_times(self._h, C.byref(f), C.byref(b)))
        return f.value, b.value

    def device_bytes(self) -> int:
        return int(_lib().hlbn_device_bytes(self._h))


# ---------------------------------------------------------------------------
# Training data: periodic states of the plain binary32 solver
# ---------------------------------------------------------------------------
# Parameter ranges of the synthetic vessels.  "slow": low viscosity, inlet
# amplitude given as a lattice density; the plain run from rest needs tens
# of cycles to become periodic.  "fast": tau 0.6-1.0 with 0.5-2 mmHg
# amplitudes, where it needs only a few.
REGIMES = {
    "slow": dict(tau=(0.52, 0.56), amp=(2e-4, 2e-3), units="lattice"),
    "fast": dict(tau=(0.6, 1.0), amp=(0.5, 2.0), units="mmHg"),
}


@dataclass
class RunSpec:
    """One vessel of the training or evaluation set."""
    kind: str            # "cylinder" (stenosed) or "bifurcation" (3 generations)
    seed: int
    tau: float
    amplitude: float     # inlet amplitude in `units`
    units: str           # "lattice" (density) or "mmHg"
    geometry: object = None
    origin: tuple = None   # crop origin in the geometry's coordinates
    params: tuple = None   # (R, length, severity, width, z0)

    @property
    def rho_amp(self) -> float:
        """Inlet density amplitude in lattice units."""
        if self.units == "lattice":
            return self.amplitude
        from .workloads import lattice_density_amplitude
        return lattice_density_amplitude(self.tau, self.amplitude)


def run_spec(seed: int, side: int = 128, regime: str = "slow") -> RunSpec:
    """Seeded vessel fitting a side^3 crop: even seeds stenosed cylinders
    (radius ~U(6, 14), length side - 8, area reduction ~U(0.2, 0.7), width
    ~U(1, 3) x radius; sizes for side 128, scaled with the crop), odd seeds
    3-generation bifurcations (root radius ~U(4.5, 6) x side / 128); tau and
    inlet amplitude uniform over the regime's ranges."""
    from .geometry import Geometry

    reg = REGIMES[regime]
    rng = np.random.default_rng(10007 * seed + 17)
    tau = float(rng.uniform(*reg["tau"]))
    amp = float(rng.uniform(*reg["amp"]))
    scale = side / 128.0
    if seed % 2 == 0:
        R = float(rng.uniform(6.0, 14.0) * scale)
        length = side - 8
        sev = float(rng.uniform(0.2, 0.7))
        width = float(rng.uniform(1.0, 3.0) * R)
        z0 = float(rng.uniform(0.3, 0.7) * length)
        g = Geometry.stenosed_cylinder(R, length, sev, z0, width)
        params = (R, length, sev, width, z0)
        kind = "cylinder"
    else:
        R0 = float(rng.uniform(4.5, 6.0) * scale)
        for attempt in range(50):
            g = Geometry.vessel_tree(seed * 1000 + attempt, 3, R0)
            c = g.links()[0]
            if ((c.max(0) - c.min(0) + 1) <= side - 4).all():
                break
        params = (R0, 0, 0.0, 0.0, 0.0)
        kind = "bifurcation"
    c = g.links()[0]
    lo, hi = c.min(0), c.max(0)
    origin = tuple(int(v) for v in lo - (side - (hi - lo + 1)) // 2)
    return RunSpec(kind, seed, tau, amp, reg["units"], g, origin, params)


def input_scalars(spec: RunSpec, period: float) -> tuple:
    """Per-run input channels 1-3: 10 (tau - 1/2), period / 1000 steps and
    half the Stokes-layer thickness sqrt(nu T / pi) in voxels."""
    nu = (spec.tau - 0.5) / 3.0
    return (10.0 * (spec.tau - 0.5), period / 1000.0, 0.5 * math.sqrt(nu * period / math.pi))


def output_scales(spec: RunSpec) -> tuple:
    """(velocity, density) units of the network outputs: the fields are
    linear in the inlet amplitude at these Reynolds numbers, so the outputs
    are u / (rho_amp / 2) and (rho - 1) / rho_amp."""
    a = spec.rho_amp
    return 0.5 * a, a


def input_field(spec: RunSpec, side: int, period: float, phase: float,
                sdf: np.ndarray | None = None) -> np.ndarray:
    """[side, side, side, 6] network input for `spec` at `phase` (radians):
    signed distance / 16 (clipped to +-1), input_scalars, cos and sin of
    the phase."""
    if sdf is None:
        sdf = spec.geometry.signed_distance(spec.origin, (side, side, side))
    x = np.empty((side, side, side, IN_CHANNELS), np.float32)
    x[..., 0] = np.clip(sdf, -16.0, 16.0) / 16.0
    for k, v in enumerate(input_scalars(spec, period)):
        x[..., 1 + k] = v
    x[..., 4] = math.cos(phase)
    x[..., 5] = math.sin(phase)
    return x


def signed_distances(specs, side: int, workers: int = 16) -> list:
    """Signed-distance crops of `specs` ([side^3] float32 each), computed in
    parallel (the native transform releases the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(s):
        return s.geometry.signed_distance(s.origin, (side, side, side)).reshape(-1)
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        return list(ex.map(one, specs))


def _crop_index(spec: RunSpec, coords: np.ndarray, offset, side: int) -> np.ndarray:
    """Flat crop index of each ensemble site of the instance."""
    local = coords - np.asarray(offset) - np.asarray(spec.origin)
    return (local[:, 0] * side + local[:, 1]) * side + local[:, 2]


def _ensemble_of(specs, period):
    from . import workloads as WL
    units = {s.units for s in specs}
    if len(units) != 1:
        raise ValueError("an ensemble's runs must share amplitude units")
    parts = [(s.geometry, *s.params, s.tau, s.amplitude) for s in specs]
    return WL.make_ensemble(parts, period, name="warm_start_batch",
                            amplitude_units=units.pop())


def run_to_periodic(lbm, ens, period: int, tol: float = 1e-4, max_cycles: int = 80,
                    start_cycle_check: int = 1):
    """Advances the ensemble cycle by cycle; returns per-instance cycles until
    max_sites |u(t) - u(t - T)| / max |u| < tol (0 where never reached, -1
    where the instance went unstable) and the number of cycles run."""
    n = len(ens.instances)
    done = np.zeros(n, np.int64)
    lbm.snapshot_velocity()
    cycles = 0
    for c in range(1, max_cycles + 1):
        lbm.step(period)
        cycles = c
        mon = lbm.monitor_instances()
        umax = np.array([m["velocity_max"] for m in mon])
        bad = ~np.array([bool(m["stable"]) for m in mon]) | ~np.isfinite(umax)
        change = lbm.velocity_change_instances(np.where(bad | (umax <= 0), 1.0, umax))
        lbm.snapshot_velocity()
        done[bad & (done == 0)] = -1
        newly = np.isfinite(change) & (change < tol) & (done == 0) & (c >= start_cycle_check)
        done[newly] = c
        if (done != 0).all():
            break
    return done, cycles


def simulate(specs, side: int, period: int, phases: int = 8, tol: float = 1e-4,
             max_cycles: int = 80, storage: str = "fp32"):
    """Runs the vessels of `specs` together from rest to their periodic state
    and samples `phases` fields over the next cycle.  Returns (targets
    [len, phases, side^3, 4] float16 in the output_scales units, masks
    [len, side^3] uint8, cycles, seconds)."""
    import time as _time
    ens = _ensemble_of(specs, period)
    lbm = ens.solver(storage=storage)
    t0 = _time.perf_counter()
    cycles, _ = run_to_periodic(lbm, ens, period, tol, max_cycles)
    n3 = side ** 3
    targets = np.zeros((len(specs), phases, n3, OUT_CHANNELS), np.float16)
    masks = np.zeros((len(specs), n3), np.uint8)
    idx = [_crop_index(s, ens.domain.coords[i.sites], i.offset, side)
           for s, i in zip(specs, ens.instances)]
    for k, ix in enumerate(idx):
        masks[k, ix] = 1
    sc = [output_scales(s) for s in specs]
    start = lbm.GetTimeStep()
    for ph in range(phases):
        target_t = start + int(round(ph * period / phases))
        lbm.step(target_t - lbm.GetTimeStep())
        rho, u = lbm.macroscopic()
        for k, (inst, ix) in enumerate(zip(ens.instances, idx)):
            sl = inst.sites
            targets[k, ph, ix, :3] = u[sl] / sc[k][0]
            targets[k, ph, ix, 3] = (rho[sl] - 1.0) / sc[k][1]
    lbm.synchronize()
    return targets, masks, cycles, _time.perf_counter() - t0


# ---------------------------------------------------------------------------
# Training and the warm-start evaluation
# ---------------------------------------------------------------------------
def _torch_cuda():
    """torch with its HIP runtime initialised.  torch ships its own HIP
    runtime next to the library's (ROCm's); in one process torch's has to
    initialise first, so every entry point here that touches torch calls
    this before the library creates a solver or a network."""
    import torch
    torch.cuda.init()
    return torch


def _device_sync():
    import torch
    torch.cuda.current_stream().synchronize()


def build_dataset(seeds, side: int, periods, batch_runs: int = 64, phases: int = 8,
                  tol: float = 1e-4, max_cycles: int = 80, regime: str = "slow", log=print):
    """Periodic-state fields of the runs `seeds`, simulated batch_runs at a
    time (each batch with its own period), kept in device memory:
    dict with sdf [R, S^3] f16, targets [R, phases, S^3, 4] f16, masks
    [R, S^3] u8, counts, and per run (spec, period, cycles)."""
    import time as _time

    torch = _torch_cuda()
    dev = torch.device("cuda")
    R = len(seeds)
    n3 = side ** 3
    sdf = torch.empty((R, n3), dtype=torch.float16, device=dev)
    tg = torch.empty((R, phases, n3, OUT_CHANNELS), dtype=torch.float16, device=dev)
    mk = torch.empty((R, n3), dtype=torch.uint8, device=dev)
    runs = []
    for b0 in range(0, R, batch_runs):
        bs = seeds[b0:b0 + batch_runs]
        period = int(periods[(b0 // batch_runs) % len(periods)])
        specs = [run_spec(s, side, regime) for s in bs]
        t0 = _time.perf_counter()
        targets, masks, cycles, secs = simulate(specs, side, period, phases, tol, max_cycles)
        for k, (s, d) in enumerate(zip(specs, signed_distances(specs, side))):
            sdf[b0 + k] = torch.from_numpy(np.clip(d, -16, 16) / 16.0).to(dev, torch.float16)
            # only runs that reached a periodic state with finite fields train
            ok = bool(cycles[k] > 0 and np.isfinite(targets[k]).all())
            if not ok:
                masks[k] = 0
            runs.append(dict(spec=s, period=period, cycles=int(cycles[k]), ok=ok))
        tg[b0:b0 + len(bs)] = torch.from_numpy(targets).to(dev)
        mk[b0:b0 + len(bs)] = torch.from_numpy(masks).to(dev)
        conv = cycles[cycles > 0]
        log(f"[warmstart] simulated runs {b0}..{b0 + len(bs) - 1} (period {period}): "
            f"{secs:.1f}s LB, {_time.perf_counter() - t0:.1f}s total, "
            f"periodic after {np.median(conv) if conv.size else 0:.0f} cycles (median, max "
            f"{conv.max() if conv.size else 0}), {(cycles == 0).sum()} not converged, "
            f"{(cycles < 0).sum()} unstable")
    return dict(sdf=sdf, targets=tg, masks=mk, counts=mk.sum(1).double().cpu().numpy(),
                runs=runs, phases=phases, side=side)


def _input_tensor(sdf, spec: RunSpec, period: float, phase: float, out=None):
    """Device input [S^3, 6] from a device signed-distance crop (already
    clipped and divided by 16)."""
    torch = _torch_cuda()
    n = sdf.numel()
    x = out if out is not None else torch.empty((n, IN_CHANNELS), dtype=torch.float32,
                                                device=sdf.device)
    x[:, 0] = sdf.reshape(-1).float()
    for k, v in enumerate(input_scalars(spec, period)):
        x[:, 1 + k] = v
    x[:, 4] = math.cos(phase)
    x[:, 5] = math.sin(phase)
    return x


def _sample_phase(ph: int, phases: int) -> float:
    # samples are taken at start + ph * T / phases, start = 1 mod T
    return 2.0 * math.pi * ph / phases


def train(net: "UNet", data, epochs: int, lr: float = 1e-3, seed: int = 0,
          channel_weights=(1.0, 1.0, 1.0, 1.0), log=print, time_budget_s: float = 1e30):
    """Adam over every (run, phase) sample, one per step, cosine-decayed
    learning rate (over the epochs or the time budget, whichever ends
    first).  Returns the per-epoch mean losses."""
    import time as _time

    torch = _torch_cuda()
    rng = np.random.default_rng(seed)
    good = np.array([r for r, run in enumerate(data["runs"]) if run["ok"]], np.int64)
    R, phases = len(good), data["phases"]
    total = epochs * R * phases
    done, losses, t0 = 0, [], _time.perf_counter()
    w = np.asarray(channel_weights, np.float32)
    x = torch.empty((data["side"] ** 3, IN_CHANNELS), dtype=torch.float32, device="cuda")
    for ep in range(epochs):
        order = rng.permutation(R * phases)
        acc, cnt = 0.0, 0
        for i in order:
            r, ph = divmod(int(i), phases)
            r = int(good[r])
            run = data["runs"][r]
            _input_tensor(data["sdf"][r], run["spec"], run["period"],
                          _sample_phase(ph, phases), out=x)
            t = data["targets"][r, ph].float()
            _device_sync()
            net.forward(x, device=True)
            loss = net.backward(t, data["masks"][r], w, float(data["counts"][r]), device=True)
            if not math.isfinite(loss):
                raise RuntimeError(f"non-finite training loss at step {done} (run {r})")
            elapsed = _time.perf_counter() - t0
            frac = min(1.0, max(done / max(total, 1), elapsed / time_budget_s))
            net.adam(lr=lr * 0.5 * (1.0 + math.cos(math.pi * frac)))
            acc += loss
            cnt += 1
            done += 1
            if done % 500 == 0:
                log(f"[warmstart] step {done}: loss {loss:.5f} ({elapsed:.0f}s)")
            if elapsed > time_budget_s:
                break
        losses.append(acc / max(cnt, 1))
        log(f"[warmstart] epoch {ep}: mean loss {losses[-1]:.5f} "
            f"({_time.perf_counter() - t0:.0f}s)")
        if _time.perf_counter() - t0 > time_budget_s:
            break
    return losses


def predict_fields(net: "UNet", specs, ens, side: int, period: int, phase: float,
                   sdfs=None):
    """Predicted density and momentum of every ensemble site as float64
    device tensors ([n], [n, 3]), the network run once per instance on the
    device."""
    torch = _torch_cuda()
    dev = torch.device("cuda")
    n = ens.domain.n
    rho = torch.ones(n, dtype=torch.float64, device=dev)
    mom = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    if sdfs is None:
        sdfs = signed_distances(specs, side)
    x = torch.empty((side ** 3, IN_CHANNELS), dtype=torch.float32, device=dev)
    y = torch.empty((side ** 3, OUT_CHANNELS), dtype=torch.float32, device=dev)
    for s, inst, d in zip(specs, ens.instances, sdfs):
        sd = torch.from_numpy(np.clip(d, -16, 16) / 16.0).to(dev, torch.float32)
        _input_tensor(sd, s, period, phase, out=x)
        _device_sync()
        net.forward(x, y, device=True)
        ix = torch.from_numpy(_crop_index(s, ens.domain.coords[inst.sites], inst.offset,
                                          side)).to(dev)
        su, sr = output_scales(s)
        yi = y.index_select(0, ix).double()
        r = 1.0 + yi[:, 3] * sr
        rho[inst.sites] = r
        mom[inst.sites] = yi[:, :3] * su * r[:, None]
    return rho, mom


def evaluate(net: "UNet", seeds, side: int, period: int, tol: float = 1e-4,
             max_cycles: int = 80, regime: str = "slow", log=print,
             start: str = "feq") -> dict:
    """Held-out vessels run to the periodic state from rest (the plain
    method) and from the network's prediction; returns the cycles and wall
    times of both and the difference of the converged fields.  start "feq":
    the equilibrium of the predicted field (on the device); "ce": with the
    Chapman-Enskog non-equilibrium part of its velocity gradient
    (chapman_enskog_f, on the host)."""
    import time as _time
    _torch_cuda()
    specs = [run_spec(s, side, regime) for s in seeds]
    ens = _ensemble_of(specs, period)
    # plain run from rest
    cold = ens.solver(storage="fp32")
    cold.synchronize()
    t0 = _time.perf_counter()
    c_cold, n_cold = run_to_periodic(cold, ens, period, tol, max_cycles)
    wall_cold = _time.perf_counter() - t0
    _, u_ref0 = cold.macroscopic()   # periodic field at the phase of step 1
    # warm start: equilibrium of the predicted field at the phase of step 1
    warm = ens.solver(storage="fp32")
    warm.synchronize()
    t0 = _time.perf_counter()
    sdfs = signed_distances(specs, side)
    t_sdf = _time.perf_counter() - t0
    rho, mom = predict_fields(net, specs, ens, side, period, 0.0, sdfs)
    if start == "ce":
        r_h = rho.cpu().numpy()
        warm.set_distributions(chapman_enskog_f(ens, specs, side, r_h,
                                                mom.cpu().numpy() / r_h[:, None]))
    else:
        warm.set_equilibrium_fields(rho, mom)
    warm.SetTimeStep(1)
    t_pred = _time.perf_counter() - t0
    c_warm, n_warm = run_to_periodic(warm, ens, period, tol, max_cycles)
    wall_warm = _time.perf_counter() - t0
    u_pred = (mom / rho[:, None]).cpu().numpy()
    pred_l2, pred_max = [], []
    for inst in ens.instances:
        a, b = u_pred[inst.sites], u_ref0[inst.sites]
        pred_l2.append(float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)))
        pred_max.append(float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)))
    # instances the plain method cannot run stably are reported, not scored
    unstable = (c_cold < 0) | (c_warm < 0)
    if unstable.all():
        raise RuntimeError("every instance went unstable")
    # converged fields at the same phase
    steps = max(n_cold, n_warm) * period
    cold.step(steps + 1 - cold.GetTimeStep())
    warm.step(steps + 1 - warm.GetTimeStep())
    _, u_c = cold.macroscopic()
    _, u_w = warm.macroscopic()
    diffs = []
    for inst, bad in zip(ens.instances, unstable):
        if bad:
            continue
        a, b = u_w[inst.sites], u_c[inst.sites]
        diffs.append(float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)))
    conv = (c_cold > 0) & (c_warm > 0)
    ratio = (c_cold[conv] / c_warm[conv]) if conv.any() else np.array([np.nan])
    res = dict(regime=regime, instances=len(specs), period=period, tol=tol,
               unstable_cold=np.flatnonzero(c_cold < 0).tolist(),
               unstable_warm=np.flatnonzero(c_warm < 0).tolist(),
               cycles_cold=c_cold.tolist(), cycles_warm=c_warm.tolist(),
               cycles_run_cold=int(n_cold), cycles_run_warm=int(n_warm),
               cycle_ratio_median=float(np.median(ratio)),
               cycle_ratio_total=float(c_cold[conv].sum() / max(c_warm[conv].sum(), 1)),
               wall_cold_s=wall_cold, wall_warm_s=wall_warm, sdf_s=t_sdf,
               prediction_s=t_pred, wall_ratio=wall_cold / max(wall_warm, 1e-9),
               prediction_rel_l2_median=float(np.median(pred_l2)),
               prediction_rel_max_median=float(np.median(pred_max)),
               converged_field_max_rel_diff=float(np.max(diffs)),
               converged_field_median_rel_diff=float(np.median(diffs)))
    log(f"[warmstart] eval: cycles until all periodic cold {n_cold} warm {n_warm}; "
        f"per-instance median ratio {res['cycle_ratio_median']:.2f}; wall {wall_cold:.2f}s "
        f"vs {wall_warm:.2f}s (x{res['wall_ratio']:.2f}, prediction {t_pred:.2f}s, "
        f"median relative L2 error {res['prediction_rel_l2_median']:.3f}); "
        f"field diff max {res['converged_field_max_rel_diff']:.2e}")
    return res


def chapman_enskog_f(ens, specs, side: int, rho: np.ndarray, u: np.ndarray) -> np.ndarray:
    """Distributions [n, 19] of a start with the first-order non-equilibrium
    part: f_i = feq_i(rho, u) - tau w_i rho / Cs2 (c_i c_i - Cs2 I) : S, with
    S the symmetric velocity gradient by central differences on each
    vessel's crop grid (solid voxels at rest) and tau the vessel's relaxation
    time (the Chapman-Enskog expansion of the lattice BGK equation)."""
    from .lb import D3Q19
    c = D3Q19.VECTORS.astype(np.float64)
    w = D3Q19.EQMWEIGHTS
    cs2 = 1.0 / 3.0
    f = np.empty((len(rho), 19))
    for s, inst in zip(specs, ens.instances):
        sl = inst.sites
        ix = _crop_index(s, ens.domain.coords[sl], inst.offset, side)
        grid = np.zeros((side ** 3, 3))
        grid[ix] = u[sl]
        grid = grid.reshape(side, side, side, 3)
        g = np.stack(np.gradient(grid, axis=(0, 1, 2)), 0)   # [a, x, y, z, b] = d_a u_b
        du = g.reshape(3, -1, 3)[:, ix, :].transpose(1, 0, 2)  # [site, a, b]
        S = 0.5 * (du + du.transpose(0, 2, 1))
        r, v = rho[sl], u[sl]
        cu = v @ c.T
        feq = w * r[:, None] * (1.0 + cu / cs2 + 0.5 * (cu / cs2) ** 2 -
                                0.5 * (v * v).sum(1)[:, None] / cs2)
        qs = np.einsum("ia,ib,sab->si", c, c, S) - cs2 * np.trace(S, axis1=1, axis2=2)[:, None]
        f[sl] = feq - s.tau * w * r[:, None] / cs2 * qs
    return f


def epsilon_study(seeds, side: int, period: int, epsilons=(0.0, 1e-3, 1e-2, 1e-1),
                  tol: float = 1e-4, max_cycles: int = 80, regime: str = "slow",
                  log=print) -> dict:
    """How many cycles a start of a given accuracy saves, without the
    network.  The plain run from rest gives every held-out vessel's periodic
    state; runs then start at the phase of step 1 from
      * "exact_f": that state's distributions themselves;
      * eps: the equilibrium of its density and velocity with relative error
        eps (rho - 1 and u scaled by 1 - eps; eps = 0 is the exact
        macroscopic field, whose start still lacks the non-equilibrium part);
      * ce_eps: the same field with the Chapman-Enskog non-equilibrium part
        of its velocity gradient (chapman_enskog_f);
    and run to the periodic state.  Returns per-vessel cycles of every
    start and the ratios to the plain run."""
    _torch_cuda()
    specs = [run_spec(s, side, regime) for s in seeds]
    ens = _ensemble_of(specs, period)
    cold = ens.solver(storage="fp32")
    c_cold, n_cold = run_to_periodic(cold, ens, period, tol, max_cycles)
    # the periodic state at the phase of step 1 (t = 1 mod period)
    cold.step((n_cold + 1) * period + 1 - cold.GetTimeStep())
    f_exact = cold.get_distributions()
    rho, u = cold.macroscopic()
    del cold
    ok = c_cold > 0
    res = dict(regime=regime, side=side, period=period, tol=tol, instances=len(specs),
               cycles_cold=c_cold.tolist(), starts={})

    def run(label, setup):
        lbm = ens.solver(storage="fp32")
        setup(lbm)
        lbm.SetTimeStep(1)
        c, _ = run_to_periodic(lbm, ens, period, tol, max_cycles)
        good = ok & (c > 0)
        ratio = c_cold[good] / c[good] if good.any() else np.array([np.nan])
        res["starts"][label] = dict(cycles=c.tolist(), ratio_median=float(np.median(ratio)),
                                    ratio_total=float(c_cold[good].sum() / max(c[good].sum(), 1)))
        log(f"[warmstart] epsilon study {label}: cycles median {np.median(c[good]):.0f} "
            f"(plain {np.median(c_cold[ok]):.0f}), per-vessel ratio median "
            f"{res['starts'][label]['ratio_median']:.2f}")

    run("exact_f", lambda lbm: lbm.set_distributions(f_exact))
    for eps in epsilons:
        r_e = 1.0 + (rho - 1.0) * (1.0 - eps)
        u_e = u * (1.0 - eps)
        m_e = u_e * r_e[:, None]
        run(f"{eps:g}", lambda lbm, r_e=r_e, m_e=m_e: lbm.set_equilibrium_fields(r_e, m_e))
        f_ce = chapman_enskog_f(ens, specs, side, r_e, u_e)
        run(f"ce_{eps:g}", lambda lbm, f_ce=f_ce: lbm.set_distributions(f_ce))
    return res


def main(argv=None) -> int:
    import argparse
    import json
    import sys
    import time as _time
    ap = argparse.ArgumentParser(description="train and evaluate the learned warm start")
    ap.add_argument("--side", type=int, default=64)
    ap.add_argument("--regime", choices=sorted(REGIMES), default="slow")
    ap.add_argument("--train-runs", type=int, default=512)
    ap.add_argument("--eval-runs", type=int, default=64)
    ap.add_argument("--batch-runs", type=int, default=64)
    ap.add_argument("--periods", default="500,400,700")
    ap.add_argument("--eval-period", type=int, default=500)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--max-cycles", type=int, default=200)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--train-seconds", type=float, default=1e30)
    ap.add_argument("--out", default="")
    ap.add_argument("--save-weights", default="")
    ap.add_argument("--load-weights", default="",
                    help="evaluate saved parameters (no data generation or training)")
    ap.add_argument("--start", choices=["feq", "ce"], default="feq",
                    help="warm start from the predicted field's equilibrium, or with its "
                         "Chapman-Enskog non-equilibrium part")
    ap.add_argument("--epsilon-study", action="store_true",
                    help="cycles saved by starts of known accuracy (no network)")
    a = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)
    _torch_cuda()
    if a.epsilon_study:
        res = epsilon_study(list(range(100000, 100000 + a.eval_runs)), a.side, a.eval_period,
                            tol=a.tol, max_cycles=a.max_cycles, regime=a.regime, log=log)
        line = json.dumps(res)
        print(line, flush=True)
        if a.out:
            open(a.out, "w").write(line + "\n")
        return 0
    net = UNet(a.side, 1)
    info = dict(side=a.side)
    if a.load_weights:
        net.set_params(np.load(a.load_weights))
        info.update(weights=a.load_weights)
    else:
        t0 = _time.perf_counter()
        periods = [int(v) for v in a.periods.split(",")]
        data = build_dataset(list(range(a.train_runs)), a.side, periods, a.batch_runs, 8,
                             a.tol, a.max_cycles, a.regime, log)
        t_data = _time.perf_counter() - t0
        net.set_params(init_params())
        t1 = _time.perf_counter()
        losses = train(net, data, a.epochs, a.lr, log=log, time_budget_s=a.train_seconds)
        t_train = _time.perf_counter() - t1
        f_ms, b_ms = net.times()
        if a.save_weights:
            np.save(a.save_weights, net.params())
        train_cycles = [r["cycles"] for r in data["runs"] if r["ok"]]
        del data
        info.update(train_runs=a.train_runs, epochs_run=len(losses), losses=losses,
                    data_seconds=t_data, train_seconds=t_train, forward_ms=f_ms,
                    backward_ms=b_ms, train_cycles_median=float(np.median(train_cycles)))
    res = evaluate(net, list(range(100000, 100000 + a.eval_runs)), a.side, a.eval_period,
                   a.tol, a.max_cycles, a.regime, log, a.start)
    res["start"] = a.start
    res.update(info)
    line = json.dumps(res)
    print(line, flush=True)
    if a.out:
        open(a.out, "w").write(line + "\n")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
