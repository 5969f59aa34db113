"""Subdomains with ghost sites and device halo plans (the reference's
decomposed run: Domain.cc:166-179 and 247-284 for the shared distributions,
FieldData.cc:27-49 for SendAndReceive / CopyReceived, lb.hpp:175-251 for
their place in the step).

A domain is cut into two halves of Morton-ordered 8^3 blocks, run as two
solvers on one GPU that exchange their halo values through device buffers
after every step, and compared bit for bit with the single-domain run (which
is itself bit-exact with the oracle)."""
import time

import numpy as np
import pytest

from hemelb_amd import lb as L
from hemelb_amd import workloads as WL
from hemelb_amd.geometry import split_domain
from tests import cases as K

pytestmark = pytest.mark.gpu


def _subdomains(traits, domain, tau, inlets, outlets, storage, rho0, layout, parts=2):
    pieces = split_domain(domain, parts)
    subs = []
    for own, _, ghosts in pieces:
        s = L.LBM(traits, own, L.LbmParameters(tau=tau), inlets, outlets, storage=storage,
                  rho0=rho0, layout=layout, ghosts=ghosts)
        s.SetInitialConditions(rho0)
        subs.append(s)
    plans = [L.HaloPlan(subs[a], subs[b]) for a in range(parts) for b in range(parts)
             if a != b]
    return pieces, subs, plans


def _compare(whole, pieces, subs, label):
    f = whole.get_distributions()
    for k, ((_, idx, _), s) in enumerate(zip(pieces, subs)):
        g = s.get_distributions()
        bad = int(np.count_nonzero(g != f[idx]))
        assert bad == 0, f"{label}: subdomain {k} differs in {bad} values"


CASES = [
    ("tree19-lbgk-bfl-ladd-wom", lambda: K.tree_case(3, 6.0, seed=1)),
    ("cyl15-lbgk-sbb-nash", lambda: K.cylinder_case(6.0, 24, 15, "LBGK", "SBB", "cosine")),
    ("cyl19-mrt-bfl-ladd-par",
     lambda: K.cylinder_case(6.0, 24, 19, "MRT", "BFL", "parabolic", tau=0.7)),
]


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16"])
@pytest.mark.parametrize("layout", ["blocks", "pull_table"])
@pytest.mark.parametrize("label,make", CASES, ids=[c[0] for c in CASES])
def test_two_subdomains_match_the_single_domain(label, make, layout, storage):
    case = make()
    print(f"[halo] {label} {layout} {storage} n={case.domain.n}", flush=True)
    whole = case.product(storage=storage, layout=layout)
    pieces, subs, plans = _subdomains(case.traits(), case.domain, case.tau,
                                      [i.to_product() for i in case.inlets],
                                      [o.to_product() for o in case.outlets], storage,
                                      case.rho0, layout)
    assert all(p.count > 0 for p in plans)
    for t in range(1, 31):
        whole.step(1)
        for s in subs:
            s.step(1)
        for p in plans:
            p.exchange()
        if t in (1, 2, 9, 30):
            _compare(whole, pieces, subs, f"{label} after {t} steps")
    assert subs[0].GetTimeStep() == whole.GetTimeStep() == 31


def test_exchange_needs_a_post_collision_state():
    case = K.tree_case(3, 6.0, seed=1)
    _, subs, plans = _subdomains(case.traits(), case.domain, case.tau,
                                 [i.to_product() for i in case.inlets],
                                 [o.to_product() for o in case.outlets], "fp64", 1.0, "blocks")
    from hemelb_amd import _native as N
    with pytest.raises(N.HemeLBError, match="post-collision"):
        plans[0].exchange()


@pytest.mark.timeout(600)
def test_vessel_tree_in_two_halves():
    """The 7.93 M-site tree in two Morton halves, 20 steps, bit-exact with
    the single-domain run; per-step pack and unpack times recorded."""
    t0 = time.time()
    w = WL.vessel_tree("LBGK")
    pieces, subs, plans = _subdomains(w.traits, w.domain, w.tau, w.inlets, w.outlets,
                                      "fp64", w.rho0, "blocks")
    print(f"[halo] tree halves: {[p[0].n for p in pieces]} sites, ghosts "
          f"{[p[2].n for p in pieces]}, exchanged values {[p.count for p in plans]} "
          f"({time.time() - t0:.1f}s)", flush=True)
    whole = w.solver(storage="fp64")
    times = []
    for t in range(1, 21):
        whole.step(1)
        for s in subs:
            s.step(1)
        for p in plans:
            p.exchange()
        times.append([p.last_times() for p in plans])
    _compare(whole, pieces, subs, "tree after 20 steps")
    ms = np.array(times[2:])  # [steps, plans, (pack, unpack)]
    print(f"[halo] per step: pack {ms[..., 0].sum(1).mean() * 1e3:.1f} us, unpack "
          f"{ms[..., 1].sum(1).mean() * 1e3:.1f} us for {sum(p.count for p in plans)} values "
          f"(both directions)", flush=True)
    assert np.isfinite(ms).all() and (ms >= 0).all()
