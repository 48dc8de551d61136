"""The device laser envelope (qz_laser_*: Laser::prepare_slice and advance_slice of every slice, src/Laser.cpp:292-419)
against the oracle's restatement (laser_oracle.inc, pinned in test_oracle_pins.py to the reference's vacuum-
diffraction validation): the envelope and the ponderomotive arrays are bit-identical after every step, for the three
xi schemes, linear and circular polarisation, in vacuum and with a slice- and node-dependent susceptibility."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import quarz_amd
    quarz_amd.lib()
    return quarz_amd


@pytest.fixture(scope="module")
def oracle():
    import oracle_lib
    return oracle_lib


@pytest.mark.parametrize("scheme", ["bdf2", "trapezoid"])
def test_laser_vacuum_validation_bit_exact(qa, oracle, scheme):
    """validation/laser_vacuum.in: the device pulse, the 40 steps and the device's Laser::sums equal the oracle's bit
    for bit, and the device run prints the reference's validation line."""
    import torch
    from laser_checks import PRINTED, VAC, diag_row, vacuum_line
    v = VAC
    nodes = oracle.make_uniform(v["R"], v["dr"])
    S = oracle.Solver(nodes)
    code = {"trapezoid": 0, "bdf2": 1}[scheme]
    L = qa.Laser(nodes, v["K"], v["xi_min"], v["dxi"], v["dt"], v["k0"], scheme=scheme)
    L.gaussian(v["a0"], v["w0"], v["L0"], v["xi0"], v["focus"])
    a = S.laser_gaussian(v["K"], v["xi_min"], v["dxi"], v["a0"], v["w0"], v["L0"], v["xi0"], v["focus"], v["k0"])
    assert np.array_equal(L.envelope(), a)
    chi_h = np.full((v["K"], S.M), v["chi"])
    chi = torch.from_numpy(chi_h).cuda()
    rows = []
    for n in range(v["rows"]):
        Sd = L.sums()
        assert np.array_equal(Sd, S.laser_sums(a, v["xi_min"], v["dxi"])), n
        rows.append(diag_row(n, v["dt"] * n, Sd))
        if n + 1 < v["rows"]:
            a, pond = S.laser_step(a, v["dxi"], v["dt"], v["k0"], code, False, chi_h)
            L.step(chi)
            assert np.array_equal(L.envelope(), a), n
            assert np.array_equal(L.ponderomotive(), pond), n
    assert vacuum_line(rows) == PRINTED[scheme]
    L.close()


@pytest.mark.parametrize("scheme", ["trapezoid", "bdf2", "euler"])
@pytest.mark.parametrize("circular", [False, True])
def test_laser_plasma_bit_exact(qa, oracle, scheme, circular):
    """A stretched grid, a susceptibility that varies by slice and node (a channel with a density ramp and noise), an
    off-focus pulse; 4 steps."""
    import torch
    nodes = oracle.make_regions(12.0, [(3.0, 0.02), (12.0, 0.08)], 1.05)
    S = oracle.Solver(nodes)
    r = np.asarray(nodes)
    K, dxi, dt, k0 = 257, 0.05, 2.5, 10.0
    rng = np.random.default_rng(11)
    ramp = np.linspace(0.5, 1.5, K)[:, None]
    chi_h = 0.2 * ramp * (1 + r[None, :] ** 2 / 16.0) * (1 + 0.05 * rng.standard_normal((K, S.M)))
    code = {"trapezoid": 0, "bdf2": 1, "euler": 2}[scheme]
    L = qa.Laser(nodes, K, 0.0, dxi, dt, k0, scheme=scheme, circular=circular)
    L.gaussian(0.5, 3.0, 2.0, 7.0, 10.0)
    a = S.laser_gaussian(K, 0.0, dxi, 0.5, 3.0, 2.0, 7.0, 10.0, k0)
    chi = torch.from_numpy(chi_h).cuda()
    for n in range(4):
        a, pond = S.laser_step(a, dxi, dt, k0, code, circular, chi_h)
        L.step(chi)
        assert np.array_equal(L.envelope(), a), n
        assert np.array_equal(L.ponderomotive(), pond), n
    L.close()


def test_laser_set_envelope_and_vacuum(qa, oracle):
    """An envelope given by the caller (random, zero at the wall), vacuum (chi = NULL), and dt = 0 (no advance:
    only the ponderomotive arrays change)."""
    nodes = oracle.make_uniform(6.0, 0.03)
    S = oracle.Solver(nodes)
    K = 97
    rng = np.random.default_rng(5)
    a0 = rng.standard_normal((K, S.M, 2)) * 0.1
    a0[:, -1, :] = 0
    L = qa.Laser(nodes, K, 0.0, 0.04, 1.0, 5.0, scheme="bdf2")
    L.set_envelope(a0)
    a, pond = S.laser_step(a0, 0.04, 1.0, 5.0, 1, False, None)
    L.step(None)
    assert np.array_equal(L.envelope(), a) and np.array_equal(L.ponderomotive(), pond)
    L.close()
    L = qa.Laser(nodes, K, 0.0, 0.04, 0.0, 5.0)
    L.set_envelope(a0)
    a, pond = S.laser_step(a0, 0.04, 0.0, 5.0, 1, False, None)
    L.step(None)
    assert np.array_equal(L.envelope(), a0) and np.array_equal(L.ponderomotive(), pond)
    L.close()


def test_laser_long_box_bit_exact_and_timed(qa, oracle):
    """The size of the long-box workload's grid (512 cells, 4001 slices; 1025 nodes of the largest grid as well):
    2 steps bit-exact; prints the device time per step (the serial march in xi sets it)."""
    import time
    import torch
    for R, N, K in ((8.0, 512, 4001), (8.0, 1024, 1001)):
        nodes = oracle.make_uniform(R, R / N)
        S = oracle.Solver(nodes)
        dxi = 0.005
        L = qa.Laser(nodes, K, 0.0, dxi, 1.0, 10.0, scheme="bdf2")
        L.gaussian(0.5, 2.0, 1.0, 10.0, 0.0)
        a = S.laser_gaussian(K, 0.0, dxi, 0.5, 2.0, 1.0, 10.0, 0.0, 10.0)
        r = np.asarray(nodes)
        chi_h = np.tile(0.1 * (1 + r ** 2 / 9.0), (K, 1))
        chi = torch.from_numpy(chi_h).cuda()
        for n in range(2):
            a, pond = S.laser_step(a, dxi, 1.0, 10.0, 1, False, chi_h)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            L.step(chi)
            ms = (time.perf_counter() - t0) * 1e3
            assert np.array_equal(L.envelope(), a), n
            assert np.array_equal(L.ponderomotive(), pond), n
        print(f"laser step M={S.M} K={K}: {ms:.2f} ms ({ms * 1e3 / K:.2f} us per slice)")
        L.close()


def test_laser_batched_equals_single(qa, oracle):
    """Four different boxes stepped together (qz_laser_step_batched): grids, box lengths, schemes, polarisations,
    a vacuum box and one that is not advanced (dt = 0) -- each equals the oracle's step of that box alone, twice."""
    import torch
    specs = [  # (nodes, K, dxi, dt, k0, scheme, circular, chi?)
        (oracle.make_uniform(14.0, 0.05), 201, 0.05, 2.0, 20.0, "bdf2", False, True),
        (oracle.make_regions(12.0, [(3.0, 0.02), (12.0, 0.08)], 1.05), 257, 0.05, 2.5, 10.0, "trapezoid", True, True),
        (oracle.make_uniform(8.0, 8.0 / 512), 129, 0.01, 1.0, 12.0, "euler", False, False),
        (oracle.make_uniform(6.0, 0.03), 97, 0.04, 0.0, 5.0, "bdf2", False, True),
    ]
    code = {"trapezoid": 0, "bdf2": 1, "euler": 2}
    rng = np.random.default_rng(21)
    boxes, lasers, chis, chi_h, a = [], [], [], [], []
    for nodes, K, dxi, dt, k0, scheme, circ, has_chi in specs:
        S = oracle.Solver(nodes)
        r = np.asarray(nodes)
        c = 0.2 * (1 + r[None, :] ** 2 / 16.0) * (1 + 0.05 * rng.standard_normal((K, S.M))) if has_chi else None
        L = qa.Laser(nodes, K, 0.0, dxi, dt, k0, scheme=scheme, circular=circ)
        a0 = S.laser_gaussian(K, 0.0, dxi, 0.3, 2.0, 1.5, K * dxi / 2, 5.0, k0)
        L.set_envelope(a0)
        boxes.append(S); lasers.append(L); chi_h.append(c); a.append(a0)
        chis.append(torch.from_numpy(c).cuda() if c is not None else None)
    for step in range(2):
        qa.laser_step_batched(lasers, chis)
        for i, (nodes, K, dxi, dt, k0, scheme, circ, _) in enumerate(specs):
            a[i], pond = boxes[i].laser_step(a[i], dxi, dt, k0, code[scheme], circ, chi_h[i])
            assert np.array_equal(lasers[i].envelope(), a[i]), (step, i)
            assert np.array_equal(lasers[i].ponderomotive(), pond), (step, i)
    for L in lasers:
        L.close()


def test_laser_batched_throughput(qa, oracle):
    """256 boxes of 4001 slices x 513 nodes in one batched step (one march workgroup per box, 46 GB of HBM): two
    boxes checked against the oracle, the time per box printed against one box stepped alone."""
    import time
    import torch
    nodes = oracle.make_uniform(8.0, 8.0 / 512)
    S = oracle.Solver(nodes)
    K, dxi, nb = 4001, 0.005, 256
    r = np.asarray(nodes)
    a0 = S.laser_gaussian(K, 0.0, dxi, 0.5, 2.0, 1.0, 10.0, 0.0, 10.0)
    chi_h = np.tile(0.1 * (1 + r ** 2 / 9.0), (K, 1))
    chi = torch.from_numpy(chi_h).cuda()
    lasers = []
    for _ in range(nb):
        L = qa.Laser(nodes, K, 0.0, dxi, 1.0, 10.0, scheme="bdf2")
        L.set_envelope(a0)
        lasers.append(L)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    qa.laser_step_batched(lasers, [chi] * nb)
    ms = (time.perf_counter() - t0) * 1e3
    a1, pond = S.laser_step(a0, dxi, 1.0, 10.0, 1, False, chi_h)
    for i in (0, nb - 1):
        assert np.array_equal(lasers[i].envelope(), a1), i
        assert np.array_equal(lasers[i].ponderomotive(), pond), i
    print(f"laser batched step: {nb} boxes of {K} x {S.M} in {ms:.1f} ms, {ms / nb:.2f} ms per box")
    for L in lasers:
        L.close()


M1_MOBILE = {"modes": 1, "xi.min": -6, "grid.dr": 0.04, "driver.n0": 0.5, "driver.x0": 0.05,
             "driver.nparticles": 40000, "driver.sigma_r": 0.5, "electrons.ppc": 1}


@pytest.mark.parametrize("deck,ov", [
    ("blowout.in", {"xi.min": -10}),                          # m = 0, one workgroup
    ("linear.in", M1_MOBILE),                                 # m = 1, one workgroup
    ("linear.in", dict(M1_MOBILE, **{"sweep.groups": 3})),    # m = 1, three cooperating workgroups
])
def test_sweep_chi_feeds_the_laser(qa, oracle, decks, deck, ov):
    """The sweep hands chi_0 of every slice to the laser march (qz_sweep_args.chi_out, qz_sim_record_chi): the
    recorded chi equals the oracle's chi_0 of each slice (src/Simulation.cpp:735), the step's fields are unchanged by
    the recording, and a laser step on the run's grid and box (xi.step 0.01) fed from the device array equals the
    oracle's laser step fed from the oracle's chi."""
    g = qa.Simulation(f"{decks}/{deck}", ov)
    o = oracle.Sim(f"{decks}/{deck}", ov)
    g.record_chi()
    oracle.lib().or_sim_record_chi(o.h, 1)
    g.step()
    o.step()
    assert np.array_equal(g.fld, o.fld)
    chi_o = o.get("chi").reshape(g.K, g.M)
    assert np.array_equal(g.chi(), chi_o)
    nodes = g.get("nodes")
    S = oracle.Solver(nodes)
    K, dxi = g.K, 0.01
    L = qa.Laser(nodes, K, 0.0, dxi, 2.0, 10.0, scheme="bdf2")
    a0 = S.laser_gaussian(K, 0.0, dxi, 0.2, 2.0, 1.0, K * dxi / 2, 0.0, 10.0)
    L.set_envelope(a0)
    L.step(g.chi_device())
    a1, pond = S.laser_step(a0, dxi, 2.0, 10.0, 1, False, chi_o)
    assert np.array_equal(L.envelope(), a1) and np.array_equal(L.ponderomotive(), pond)
    L.close()
    g.record_chi(False)
    assert g.chi_device() == 0
