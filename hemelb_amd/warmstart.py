"""Learned warm start: a 3D U-Net, trained and run by the library's HIP
kernels (``include/hemelb_unet.h``), that maps a signed-distance wall field
plus iolet-parameter channels to the periodic-state velocity and pressure
of a vessel; the LB run then starts from the equilibrium of that field.

Input channels (per voxel of a cubic crop): signed distance to the wall in
voxels (positive in the fluid) divided by 16, inlet pressure amplitude
(mmHg), period / 1000 steps, tau, and cos / sin of the phase of the cycle.
Output channels: the velocity in units of 0.01 and (density - 1) / 0.01.
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
U_SCALE = 0.01
RHO_SCALE = 0.01

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
@dataclass
class RunSpec:
    """One vessel of the training or evaluation set."""
    kind: str            # "cylinder" (stenosed) or "bifurcation" (3 generations)
    seed: int
    tau: float
    amplitude_mmHg: float
    geometry: object = None
    origin: tuple = None   # crop origin in the geometry's coordinates
    params: tuple = None   # (R, length, severity, width, z0)


def run_spec(seed: int, side: int = 128) -> RunSpec:
    """Seeded vessel fitting a side^3 crop: even seeds stenosed cylinders
    (radius ~U(6, 14), length side - 8, area reduction ~U(0.2, 0.7), width
    ~U(1, 3) x radius), odd seeds 3-generation bifurcations (root radius
    ~U(4.5, 6) scaled with the crop); tau ~U(0.6, 1.0), inlet amplitude
    ~U(0.5, 2) mmHg."""
    from .geometry import Geometry

    rng = np.random.default_rng(10007 * seed + 17)
    tau = float(rng.uniform(0.6, 1.0))
    amp = float(rng.uniform(0.5, 2.0))
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
    return RunSpec(kind, seed, tau, amp, g, origin, params)


def input_field(spec: RunSpec, side: int, period: float, phase: float,
                sdf: np.ndarray | None = None) -> np.ndarray:
    """[side, side, side, 6] network input for `spec` at `phase` (radians)."""
    if sdf is None:
        sdf = spec.geometry.signed_distance(spec.origin, (side, side, side))
    x = np.empty((side, side, side, IN_CHANNELS), np.float32)
    x[..., 0] = np.clip(sdf, -16.0, 16.0) / 16.0
    x[..., 1] = spec.amplitude_mmHg
    x[..., 2] = period / 1000.0
    x[..., 3] = spec.tau
    x[..., 4] = math.cos(phase)
    x[..., 5] = math.sin(phase)
    return x


def _crop_index(spec: RunSpec, coords: np.ndarray, offset, side: int) -> np.ndarray:
    """Flat crop index of each ensemble site of the instance."""
    local = coords - np.asarray(offset) - np.asarray(spec.origin) + 0
    return (local[:, 0] * side + local[:, 1]) * side + local[:, 2]


def _ensemble_of(specs, period):
    from . import workloads as WL
    parts = [(s.geometry, *s.params, s.tau, s.amplitude_mmHg) for s in specs]
    return WL.make_ensemble(parts, period, name="warm_start_batch")


def run_to_periodic(lbm, ens, period: int, tol: float = 1e-4, max_cycles: int = 80,
                    start_cycle_check: int = 1):
    """Advances the ensemble cycle by cycle; returns per-instance cycles until
    max_sites |u(t) - u(t - T)| / max |u| < tol (0 where never reached) and
    the number of cycles run."""
    n = len(ens.instances)
    done = np.zeros(n, np.int64)
    lbm.snapshot_velocity()
    cycles = 0
    for c in range(1, max_cycles + 1):
        lbm.step(period)
        cycles = c
        umax = np.array([max(m["velocity_max"], 1e-12) for m in lbm.monitor_instances()])
        change = lbm.velocity_change_instances(umax)
        lbm.snapshot_velocity()
        newly = (change < tol) & (done == 0) & (c >= start_cycle_check)
        done[newly] = c
        if (done > 0).all():
            break
    return done, cycles


def simulate(specs, side: int, period: int, phases: int = 8, tol: float = 1e-4,
             max_cycles: int = 80, storage: str = "fp32"):
    """Runs the vessels of `specs` together from rest to their periodic state
    and samples `phases` fields over the next cycle.  Returns (targets
    [len, phases, side^3, 4] float16, masks [len, side^3] uint8, cycles,
    seconds)."""
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
    start = lbm.GetTimeStep()
    for ph in range(phases):
        target_t = start + int(round(ph * period / phases))
        lbm.step(target_t - lbm.GetTimeStep())
        rho, u = lbm.macroscopic()
        for k, (inst, ix) in enumerate(zip(ens.instances, idx)):
            sl = inst.sites
            targets[k, ph, ix, :3] = u[sl] / U_SCALE
            targets[k, ph, ix, 3] = (rho[sl] - 1.0) / RHO_SCALE
    lbm.synchronize()
    return targets, masks, cycles, _time.perf_counter() - t0


def phase_of(ph: int, phases: int, start_step: int, period: int) -> float:
    """Inlet phase (radians) of sample ph: the cosine iolet uses the 0-indexed
    step t - 1 (InOutLetCosine.cc:26-43)."""
    t = start_step + int(round(ph * period / phases)) - 1
    return 2.0 * math.pi * (t % period) / period


def equilibrium(rho: np.ndarray, u: np.ndarray, Q: int = 19) -> np.ndarray:
    """f_eq(rho, rho u) per site ([n, Q]), the textbook compressible form."""
    from . import lb as L
    lat = L.LATTICES[Q]
    c = np.asarray(lat.VECTORS, np.float64)
    w = np.asarray(lat.EQMWEIGHTS, np.float64)
    cu = u @ c.T
    uu = (u * u).sum(1)[:, None]
    return rho[:, None] * w * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu)


# ---------------------------------------------------------------------------
# Training and the warm-start evaluation
# ---------------------------------------------------------------------------
def _device_sync():
    import torch
    torch.cuda.current_stream().synchronize()


def build_dataset(seeds, side: int, periods, batch_runs: int = 64, phases: int = 8,
                  tol: float = 1e-4, max_cycles: int = 80, log=print):
    """Periodic-state fields of the runs `seeds`, simulated batch_runs at a
    time (each batch with its own period), kept in device memory:
    dict with sdf [R, S^3] f16, targets [R, phases, S^3, 4] f16, masks
    [R, S^3] u8, counts, and per run (spec, period, start step, cycles)."""
    import time as _time

    import torch
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
        specs = [run_spec(s, side) for s in bs]
        t0 = _time.perf_counter()
        targets, masks, cycles, secs = simulate(specs, side, period, phases, tol, max_cycles)
        for k, s in enumerate(specs):
            d = s.geometry.signed_distance(s.origin, (side, side, side)).reshape(-1)
            sdf[b0 + k] = torch.from_numpy(np.clip(d, -16, 16) / 16.0).to(dev, torch.float16)
            runs.append(dict(spec=s, period=period, cycles=int(cycles[k])))
        tg[b0:b0 + len(bs)] = torch.from_numpy(targets).to(dev)
        mk[b0:b0 + len(bs)] = torch.from_numpy(masks).to(dev)
        log(f"[warmstart] simulated runs {b0}..{b0 + len(bs) - 1} (period {period}): "
            f"{secs:.1f}s LB, {_time.perf_counter() - t0:.1f}s total, "
            f"periodic after {np.median(cycles):.0f} cycles (median), "
            f"{(cycles == 0).sum()} not converged")
    return dict(sdf=sdf, targets=tg, masks=mk, counts=mk.sum(1).double().cpu().numpy(),
                runs=runs, phases=phases, side=side)


def _input_tensor(data, r: int, phase: float):
    import torch
    run = data["runs"][r]
    s = data["side"]
    x = torch.empty((s ** 3, IN_CHANNELS), dtype=torch.float32, device="cuda")
    x[:, 0] = data["sdf"][r].float()
    x[:, 1] = run["spec"].amplitude_mmHg
    x[:, 2] = run["period"] / 1000.0
    x[:, 3] = run["spec"].tau
    x[:, 4] = math.cos(phase)
    x[:, 5] = math.sin(phase)
    return x


def _sample_phase(ph: int, phases: int) -> float:
    # samples are taken at start + ph * T / phases, start = 1 mod T
    return 2.0 * math.pi * ph / phases


def train(net: "UNet", data, epochs: int, lr: float = 1e-3, seed: int = 0,
          channel_weights=(1.0, 1.0, 1.0, 1.0), log=print, time_budget_s: float = 1e30):
    """Adam over every (run, phase) sample, one per step, cosine-decayed
    learning rate.  Returns the per-epoch mean losses."""
    import time as _time

    import torch
    rng = np.random.default_rng(seed)
    R, phases = len(data["runs"]), data["phases"]
    total = epochs * R * phases
    done, losses, t0 = 0, [], _time.perf_counter()
    w = np.asarray(channel_weights, np.float32)
    for ep in range(epochs):
        order = rng.permutation(R * phases)
        acc = 0.0
        for i in order:
            r, ph = divmod(int(i), phases)
            x = _input_tensor(data, r, _sample_phase(ph, phases))
            t = data["targets"][r, ph].float()
            _device_sync()
            net.forward(x, device=True)
            loss = net.backward(t, data["masks"][r], w, float(data["counts"][r]), device=True)
            elapsed = _time.perf_counter() - t0
            frac = min(1.0, max(done / max(total, 1), elapsed / time_budget_s))
            net.adam(lr=lr * 0.5 * (1.0 + math.cos(math.pi * frac)))
            acc += loss
            done += 1
            if done % 500 == 0:
                log(f"[warmstart] step {done}: loss {loss:.5f} ({elapsed:.0f}s)")
            if elapsed > time_budget_s:
                break
        losses.append(acc / max(done - ep * R * phases, 1))
        log(f"[warmstart] epoch {ep}: mean loss {losses[-1]:.5f} "
            f"({_time.perf_counter() - t0:.0f}s)")
        if _time.perf_counter() - t0 > time_budget_s:
            break
    return losses


def predict_sites(net: "UNet", spec: RunSpec, side: int, period: int, phase: float,
                  coords: np.ndarray, offset) -> tuple:
    """Predicted (rho, u) at the ensemble sites `coords` of the instance."""
    x = input_field(spec, side, period, phase)[None]
    y = net.forward(x)[0].reshape(-1, OUT_CHANNELS)
    ix = _crop_index(spec, coords, offset, side)
    u = y[ix, :3].astype(np.float64) * U_SCALE
    rho = 1.0 + y[ix, 3].astype(np.float64) * RHO_SCALE
    return rho, u


def evaluate(net: "UNet", seeds, side: int, period: int, tol: float = 1e-4,
             max_cycles: int = 80, log=print) -> dict:
    """Held-out vessels run to the periodic state from rest (the plain
    method) and from the network's prediction; returns the cycles and wall
    times of both and the difference of the converged fields."""
    import time as _time
    specs = [run_spec(s, side) for s in seeds]
    ens = _ensemble_of(specs, period)
    # plain run from rest
    cold = ens.solver(storage="fp32")
    t0 = _time.perf_counter()
    c_cold, n_cold = run_to_periodic(cold, ens, period, tol, max_cycles)
    wall_cold = _time.perf_counter() - t0
    # warm start: equilibrium of the predicted field at the phase of step 1
    warm = ens.solver(storage="fp32")
    t0 = _time.perf_counter()
    f = np.empty((ens.domain.n, 19))
    phase0 = 0.0
    for s, inst in zip(specs, ens.instances):
        rho, u = predict_sites(net, s, side, period, phase0, ens.domain.coords[inst.sites],
                               inst.offset)
        f[inst.sites] = equilibrium(rho, u)
    warm.set_distributions(f)
    warm.SetTimeStep(1)
    t_pred = _time.perf_counter() - t0
    c_warm, n_warm = run_to_periodic(warm, ens, period, tol, max_cycles)
    wall_warm = _time.perf_counter() - t0
    # converged fields at the same phase
    steps = max(n_cold, n_warm) * period
    cold.step(steps + 1 - cold.GetTimeStep())
    warm.step(steps + 1 - warm.GetTimeStep())
    _, u_c = cold.macroscopic()
    _, u_w = warm.macroscopic()
    diffs = []
    for inst in ens.instances:
        a, b = u_w[inst.sites], u_c[inst.sites]
        diffs.append(float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)))
    conv = (c_cold > 0) & (c_warm > 0)
    ratio = (c_cold[conv] / c_warm[conv]) if conv.any() else np.array([np.nan])
    res = dict(instances=len(specs), period=period, tol=tol,
               cycles_cold=c_cold.tolist(), cycles_warm=c_warm.tolist(),
               cycle_ratio_median=float(np.median(ratio)),
               cycle_ratio_total=float(c_cold[conv].sum() / max(c_warm[conv].sum(), 1)),
               wall_cold_s=wall_cold, wall_warm_s=wall_warm, prediction_s=t_pred,
               wall_ratio=wall_cold / max(wall_warm, 1e-9),
               converged_field_max_rel_diff=float(np.max(diffs)),
               converged_field_median_rel_diff=float(np.median(diffs)))
    log(f"[warmstart] eval: cycles cold {np.sum(c_cold)} warm {np.sum(c_warm)} "
        f"(median ratio {res['cycle_ratio_median']:.2f}); wall {wall_cold:.1f}s vs "
        f"{wall_warm:.1f}s (x{res['wall_ratio']:.2f}); field diff max "
        f"{res['converged_field_max_rel_diff']:.2e}")
    return res


def main(argv=None) -> int:
    import argparse
    import json
    import sys
    import time as _time
    ap = argparse.ArgumentParser(description="train and evaluate the learned warm start")
    ap.add_argument("--side", type=int, default=128)
    ap.add_argument("--train-runs", type=int, default=512)
    ap.add_argument("--eval-runs", type=int, default=64)
    ap.add_argument("--batch-runs", type=int, default=64)
    ap.add_argument("--periods", default="1000,700,1400")
    ap.add_argument("--eval-period", type=int, default=1000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--max-cycles", type=int, default=60)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--train-seconds", type=float, default=1e30)
    ap.add_argument("--out", default="")
    ap.add_argument("--save-weights", default="")
    a = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)
    t0 = _time.perf_counter()
    periods = [int(v) for v in a.periods.split(",")]
    data = build_dataset(list(range(a.train_runs)), a.side, periods, a.batch_runs, 8, a.tol,
                         a.max_cycles, log)
    t_data = _time.perf_counter() - t0
    net = UNet(a.side, 1)
    net.set_params(init_params())
    t1 = _time.perf_counter()
    losses = train(net, data, a.epochs, a.lr, log=log, time_budget_s=a.train_seconds)
    t_train = _time.perf_counter() - t1
    f_ms, b_ms = net.times()
    if a.save_weights:
        np.save(a.save_weights, net.params())
    del data
    res = evaluate(net, list(range(100000, 100000 + a.eval_runs)), a.side, a.eval_period,
                   a.tol, a.max_cycles, log)
    res.update(side=a.side, train_runs=a.train_runs, epochs_run=len(losses), losses=losses,
               data_seconds=t_data, train_seconds=t_train, forward_ms=f_ms, backward_ms=b_ms)
    line = json.dumps(res)
    print(line, flush=True)
    if a.out:
        open(a.out, "w").write(line + "\n")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
