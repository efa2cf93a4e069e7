"""The hand-written adjoint (taichimps_amd/adjoint.py, csrc/adjoint.hip)
against central finite differences.

The loss is the final mean pressure (the deform/pressure servo's pressure,
virial plus kinetic) of an isotropic-compression run. Gradients with respect
to kn, kt, gamman, gammat, xmu and to per-particle radii are compared with:
  - central differences of the FP64 device forward (relative step 1e-6);
  - central differences of the CPU oracle's forward, an independent FP64
    restatement (the FP64 reference gradient).
The small case is dense and fast (many contacts, spins, Coulomb slip), with a
small skin so the run rebuilds its list several times, and it spans several
checkpoint segments.
"""

import numpy as np
import pytest

import simhelpers as H
from oracle import oracle as O
from taichimps_amd import adjoint as ADJ
from taichimps_amd import workloads

pytestmark = pytest.mark.gpu

COEFFS0 = np.array([2.0e5, 2.0 / 7.0 * 2.0e5, 50.0, 25.0, 0.5])
MEAN_P = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0]) / 3.0


def small_case(seed=31, n=6):
    w = workloads.lattice(n, n, n, 1.0, seed, "uniform")
    rng = np.random.default_rng(seed + 1)
    w["v"] = rng.normal(0.0, 3.0, size=w["v"].shape)
    w["omega"] = rng.normal(0.0, 5.0, size=w["v"].shape)
    return w


def device_pmean(build, coeffs, radius, nsteps, every):
    sim = build(coeffs, radius)
    t, z = ADJ.run_forward(sim, nsteps, every)
    return float(MEAN_P @ t), sim, z


def central(f, x, idx, rel=1e-6):
    h = rel * abs(x[idx])
    xp, xm = x.copy(), x.copy()
    xp[idx] += h
    xm[idx] -= h
    return (f(xp) - f(xm)) / (2.0 * h)


def assert_close(adj, fd, tol, what):
    scale = max(abs(fd), 1e-12)
    err = abs(adj - fd) / scale
    assert err < tol, f"{what}: adjoint {adj:.10e} vs finite difference {fd:.10e} (rel {err:.2e})"


def assert_gradient(adj, fd, theta, tol, what):
    """Per component where the sensitivity theta_k dP/dtheta_k is resolvable
    (at least 1e-3 of the largest), and for the whole log-scaled vector."""
    s_adj, s_fd = np.abs(theta) * adj, np.abs(theta) * fd
    big = np.abs(s_fd).max()
    vec = np.abs(s_adj - s_fd).max() / big
    assert vec < tol, f"{what}: log-scaled gradient differs by {vec:.2e} of its largest entry"
    for k in range(len(adj)):
        if abs(s_fd[k]) >= 1e-3 * big:
            assert_close(adj[k], fd[k], tol, f"{what}[{k}]")


def persistent_case(seed=61, n=6):
    """A compressed lattice whose contacts neither open nor close during the
    run (spacing 0.95, radii U[0.49, 0.51], small jitter and velocities),
    with spins that keep contacts sliding when friction is low. Finite
    differences are smooth here: contact onset is where the hooke force jumps
    and the hertz force has an infinite slope, so runs in which the
    perturbation moves an onset across a step boundary are no test of a
    derivative."""
    a = 0.95
    w = workloads.lattice(n, n, n, a, seed, "uniform")
    rng = np.random.default_rng(seed + 1)
    ix, iy, iz = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    sites = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1) * a + 0.5 * a
    w["x"] = sites + rng.uniform(-0.005, 0.005, size=sites.shape)
    w["radius"] = rng.uniform(0.49, 0.51, size=len(sites))
    w["v"] = rng.normal(0.0, 0.01, size=sites.shape)
    w["omega"] = rng.normal(0.0, 2.0, size=sites.shape)
    return w


@pytest.mark.parametrize("style", ["hertz/history", "hooke/history"])
@pytest.mark.parametrize("xmu", [0.05, 50.0])
def test_adjoint_matches_finite_differences(style, xmu):
    """Persistent contacts; xmu 0.05: contacts slide (the Coulomb branch and
    its history rescale), xmu 50: every contact sticks. The servo expands the
    box and the small skin makes the run rebuild its list."""
    w = persistent_case()
    skin, nsteps, every = 0.004, 60, 25
    build = ADJ.isotropic_builder(w, style=style, skin=skin, p_gain=2e-4)
    radius = w["radius"].copy()
    coeffs = COEFFS0.copy()
    coeffs[4] = xmu
    p0, sim, z = device_pmean(build, coeffs, radius, nsteps, every)
    assert z > 4.0, "the lattice should stay compressed"
    gc, gr, _ = ADJ.run_backward(sim, MEAN_P)
    assert np.all(np.isfinite(gc)) and np.all(np.isfinite(gr))
    fd = np.array([central(lambda c: device_pmean(build, c, radius, nsteps, every)[0], coeffs, k,
                           rel=1e-6) for k in range(5)])
    assert_gradient(gc, fd, coeffs, 1e-4, f"{style} xmu {xmu} d<P>/d(kn kt gamman gammat xmu)")
    for i in (0, 37, 101, 215):
        fdr = central(lambda r: device_pmean(build, coeffs, r, nsteps, every)[0], radius, i,
                      rel=1e-6)
        assert_close(gr[i], fdr, 1e-4, f"{style} d<P>/dr[{i}]")


def test_adjoint_in_a_colliding_packing():
    """A loose, fast packing: contacts form and break, the list rebuilds
    often, many contacts slip (gran/hertz/history, continuous at onset)."""
    w = small_case()
    skin, nsteps, every = 0.02, 60, 25
    build = ADJ.isotropic_builder(w, style="hertz/history", skin=skin, p_gain=2e-4)
    radius = w["radius"].copy()
    coeffs = COEFFS0.copy()
    coeffs[4] = 0.2
    p0, sim, z = device_pmean(build, coeffs, radius, nsteps, every)
    assert sim.neighbor.build_count > 2, "the case should rebuild its list"
    gc, gr, _ = ADJ.run_backward(sim, MEAN_P)
    fd = np.array([central(lambda c: device_pmean(build, c, radius, nsteps, every)[0], coeffs, k,
                           rel=1e-6) for k in range(5)])
    assert_gradient(gc, fd, coeffs, 1e-4, "colliding d<P>/d(kn kt gamman gammat xmu)")


def test_adjoint_matches_the_oracle_fp64_reference_gradient():
    """Finite differences of the CPU oracle's forward (independent of the
    device code) against the device adjoint."""
    style, skin, nsteps, every = "hertz/history", 0.02, 60, 20
    w = small_case(seed=41)
    build = ADJ.isotropic_builder(w, style=style, skin=skin, p_gain=2e-4)
    radius = w["radius"].copy()
    coeffs = COEFFS0.copy()
    coeffs[4] = 0.2
    p_dev, sim, _ = device_pmean(build, coeffs, radius, nsteps, every)
    gc, gr, _ = ADJ.run_backward(sim, MEAN_P)

    def pm(c, r):
        w2 = dict(w)
        w2["radius"] = np.asarray(r, dtype=np.float64)
        o = O.from_workload(w2)
        o.set_neigh(skin)
        kn, kt, gn, gt, xmu = c
        o.set_pair(O.legacy_pair(style, kn=kn, gamman=gn, kt=kt, gammat=gt, xmu=xmu))
        o.set_servo(O.servo_params(p_target=1.0, p_gain=2e-4))
        assert o.run(nsteps, 1e-4) == 0
        return float(MEAN_P @ o.last_pressure_tensor())

    assert pm(coeffs, radius) == p_dev  # the FP64 forward is bitwise the oracle's
    fd = np.array([central(lambda c: pm(c, radius), coeffs, k, rel=1e-6) for k in range(5)])
    assert_gradient(gc, fd, coeffs, 1e-4, "oracle d<P>/d(kn kt gamman gammat xmu)")
    for i in (3, 150):
        fdr = central(lambda r: pm(coeffs, r), radius, i, rel=1e-6)
        assert_close(gr[i], fdr, 1e-4, f"oracle d<P>/dr[{i}]")


def test_checkpoint_interval_does_not_change_the_gradient():
    w = small_case(seed=43)
    build = ADJ.isotropic_builder(w, skin=0.02)
    radius = w["radius"].copy()
    out = []
    for every in (7, 50):
        _, sim, _ = device_pmean(build, COEFFS0, radius, 50, every)
        gc, gr, _ = ADJ.run_backward(sim, MEAN_P)
        out.append((gc, gr))
    assert np.allclose(out[0][0], out[1][0], rtol=1e-12, atol=0.0)
    assert np.allclose(out[0][1], out[1][1], rtol=1e-12, atol=1e-300)


def test_backward_leaves_the_forward_end_state():
    w = small_case(seed=47)
    build = ADJ.isotropic_builder(w, skin=0.02)
    _, sim, _ = device_pmean(build, COEFFS0, w["radius"], 30, 10)
    before = H.device_state(sim)
    ADJ.run_backward(sim, MEAN_P)
    after = H.device_state(sim)
    for k in ("x", "v", "omega", "f", "boxlo", "boxhi"):
        assert np.array_equal(before[k], after[k]), k
    # the run continues from there
    sim.run(5)


def test_torch_function_chains_the_loss():
    torch = pytest.importorskip("torch")
    w = small_case(seed=53)
    build = ADJ.isotropic_builder(w, skin=0.02)
    coeffs = torch.tensor(COEFFS0, dtype=torch.float64, requires_grad=True)
    radius = torch.tensor(w["radius"], dtype=torch.float64, requires_grad=True)
    P, Z = ADJ.final_pressure(coeffs, radius, build, 40, 15)
    assert not Z.requires_grad
    target = 1.25
    loss = (P[:3].mean() - target) ** 2
    loss.backward()
    p = float(P[:3].mean())
    _, sim, _ = device_pmean(build, COEFFS0, w["radius"], 40, 15)
    gc, gr, _ = ADJ.run_backward(sim, MEAN_P)
    assert np.allclose(coeffs.grad.numpy(), 2.0 * (p - target) * gc, rtol=1e-12)
    assert np.allclose(radius.grad.numpy(), 2.0 * (p - target) * gr, rtol=1e-12)


@pytest.mark.parametrize("style,xmu", [("hertz/history", 0.5), ("hooke/history", 0.5)])
def test_adjoint_equals_the_tangent_linear_run(style, xmu):
    """A check the collision sequence cannot spoil: the tangent-linear
    restatement (oracle/tlm.cpp) differentiates the same discrete run forward,
    with the forward's branches (contact onsets, slips, rebuilds), and no step
    size. On the adjoint case's recipe (16^3, 600 steps, rebuilds, contacts
    forming and breaking) the device forward is bitwise the restatement's FP64
    run, and the adjoint's d<P>/d(kn, kt, gamman, gammat, xmu) and its
    derivative along a random radius direction equal the tangent-linear ones
    to 1e-8 of the largest sensitivity and 1e-6 of each resolvable one."""
    c = workloads.ISOTROPIC
    w = workloads.case("adjoint_32k", dims=(16, 16, 16))
    steps, every = 600, 100
    build = ADJ.isotropic_builder(w, style=style, skin=c["skin"], dt=c["dt"],
                                  p_target=c["p_target"], p_gain=c["p_gain"])
    radius = np.asarray(w["radius"], dtype=np.float64).copy()
    coeffs = np.array([c["kn"], c["kt"], c["gamman"], c["gammat"], xmu])
    p_dev, sim, _ = device_pmean(build, coeffs, radius, steps, every)
    builds = sim.neighbor.build_count
    gc, gr, _ = ADJ.run_backward(sim, MEAN_P)
    u = np.random.default_rng(9).normal(size=len(radius))
    t, d, tl_builds = O.tlm_tangent(w, coeffs, steps, dradius=u, style=style, skin=c["skin"],
                                    dt=c["dt"], p_target=c["p_target"], p_gain=c["p_gain"])
    assert float(MEAN_P @ t) == p_dev and tl_builds == builds > 2
    tl = MEAN_P @ d
    s_adj = np.append(coeffs * gc, float(u @ gr))
    s_tl = np.append(coeffs * tl[:5], tl[5])
    big = np.abs(s_tl).max()
    print(f"{style}: adjoint {s_adj}, tangent {s_tl}")
    assert np.abs(s_adj - s_tl).max() <= 1e-8 * big, (s_adj, s_tl)
    for k in range(6):
        if abs(s_tl[k]) >= 1e-3 * big:
            assert abs(s_adj[k] - s_tl[k]) <= 1e-6 * abs(s_tl[k]), (k, s_adj[k], s_tl[k])
