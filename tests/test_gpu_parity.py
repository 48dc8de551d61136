"""GPU parity: the gfx950 path (libquarz_amd.so, through its C-ABI) against the oracle on the same inputs.

The bar is bit-identity: every array the reference's serial build would produce (fields, diagnostics, beam
sources, particles) must be equal element for element (np.array_equal; -0.0 == +0.0).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def qa():
    import quarz_amd
    quarz_amd.lib()
    return quarz_amd


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b)


def _report(name, a, b):
    a, b = np.asarray(a), np.asarray(b)
    bad = np.nonzero(a != b)
    n = len(bad[0])
    first = tuple(int(x[0]) for x in bad) if n else None
    return f"{name}: {n} of {a.size} differ; first at {first}: {a[first] if n else ''} vs {b[first] if n else ''}"


# ------------------------------------------------------------------------------------------ Thomas
@pytest.mark.parametrize("M", [16, 100, 513, 1537])
def test_thomas_batched_bit_exact(torch_dev, qa, oracle, M):
    torch = torch_dev
    rng = np.random.default_rng(M)
    batch = 7
    a = rng.uniform(-1, 1, (batch, M)); a[:, 0] = 0
    c = rng.uniform(-1, 1, (batch, M)); c[:, -1] = 0
    b = 2.5 + rng.uniform(-1, 1, (batch, M))
    d = rng.uniform(-1, 1, (batch, M))
    dev = [torch.tensor(v, dtype=torch.float64, device="cuda") for v in (a, b, c, d)]
    x = torch.zeros((batch, M), dtype=torch.float64, device="cuda")
    qa.thomas_batched(*dev, x)
    # the same systems element-major (element i of system s at i * batch + s)
    devT = [t.t().contiguous() for t in dev]
    xT = torch.zeros((M, batch), dtype=torch.float64, device="cuda")
    qa.thomas_batched(*devT, xT, element_major=True)
    torch.cuda.synchronize()
    got = x.cpu().numpy()
    gotT = xT.cpu().numpy().T
    for s in range(batch):
        ref = oracle.thomas(a[s].copy(), b[s].copy(), c[s].copy(), d[s].copy())
        assert _same(got[s], ref), _report(f"system {s}", got[s], ref)
        assert _same(gotT[s], ref), _report(f"element-major system {s}", gotT[s], ref)


# ------------------------------------------------------------------------------------------ beam kernels
def _grid_dev(torch, s):
    V, iV, ih = s.grid_arrays()
    r = torch.tensor(s.nodes, dtype=torch.float64, device="cuda")
    return (r, torch.tensor(ih, dtype=torch.float64, device="cuda"), torch.tensor(iV, dtype=torch.float64, device="cuda"))


def _beam_parts(rng, n, R, xi_lo, xi_hi, m1, gamma=2000.0):
    r = R * np.sqrt(rng.uniform(0, 1.02, n))
    th = rng.uniform(0, 2 * np.pi, n) if m1 else np.zeros(n)
    parts = {
        "x": r * np.cos(th), "y": r * np.sin(th),
        "px": rng.normal(0, 2.0, n), "py": rng.normal(0, 2.0, n), "pz": gamma * (1 + 0.01 * rng.normal(0, 1, n)),
        "xi": rng.uniform(xi_lo, xi_hi, n), "w": rng.uniform(0.5, 1.5, n) * 1e-4,
    }
    parts["w"][rng.uniform(0, 1, n) < 0.02] = 0.0
    return parts


@pytest.mark.parametrize("m1", [False, True])
@pytest.mark.parametrize("ngp", [False, True])
def test_beam_deposit_bit_exact(torch_dev, qa, oracle, m1, ngp):
    torch = torch_dev
    rng = np.random.default_rng(11 + m1 + 2 * ngp)
    nodes = oracle.make_regions(4.0, [(0.3, 0.005), (4.0, 0.05)], 1.05)
    s = oracle.Solver(nodes)
    nxi, dxi, xi_min = 120, 0.02, 0.0
    n = 60000
    parts = _beam_parts(rng, n, 3.0, -0.1, nxi * dxi + 0.1, m1)
    # a dense core: many particles per cell
    parts["x"][:5000] *= 0.01
    parts["y"][:5000] *= 0.01
    nc = 15 if m1 else 5
    ref = oracle.beam_deposit(s, xi_min, dxi, nxi, ngp, m1, -1.0, parts, nc)
    r, ih, iV = _grid_dev(torch, s)
    g = qa.radial_grid(r, ih, iV)
    box = qa.BeamGridC(xi_min, dxi, nxi, 0, nxi, int(ngp))
    p = {k: torch.tensor(parts[k], dtype=torch.float64, device="cuda") for k in qa.BEAM_ARRAYS}
    bsrc = torch.zeros((nxi, s.M, nc), dtype=torch.float64, device="cuda")
    work = torch.empty(qa.beam_deposit_workspace_bytes(n, nxi, s.M - 1), dtype=torch.uint8, device="cuda")
    qa.beam_deposit(g, box, 1 if m1 else 0, -1.0, p, bsrc, work)
    torch.cuda.synchronize()
    got = bsrc.cpu().numpy()
    assert _same(got, ref), _report("bsrc", got, ref)


@pytest.mark.parametrize("m1", [False, True])
def test_beam_deposit_beams_equals_one_after_another(torch_dev, qa, oracle, m1):
    """Two beams of different charge in one deposit pass: bit-identical to the oracle depositing the first
    beam and then the second into the same array (a row window of one pass included)."""
    torch = torch_dev
    rng = np.random.default_rng(31 + m1)
    nodes = oracle.make_regions(4.0, [(0.3, 0.005), (4.0, 0.05)], 1.05)
    s = oracle.Solver(nodes)
    nxi, dxi, xi_min = 100, 0.02, 0.0
    A = _beam_parts(rng, 40000, 3.0, -0.1, nxi * dxi + 0.1, m1)
    B = _beam_parts(rng, 15000, 1.0, 0.5, 1.5, m1)
    nc = 15 if m1 else 5
    ref = oracle.beam_deposit(s, xi_min, dxi, nxi, False, m1, -1.0, A, nc)
    ref = oracle.beam_deposit(s, xi_min, dxi, nxi, False, m1, 2.0, B, nc, into=ref)
    r, ih, iV = _grid_dev(torch, s)
    g = qa.radial_grid(r, ih, iV)
    pa = {k: torch.tensor(A[k], dtype=torch.float64, device="cuda") for k in qa.BEAM_ARRAYS}
    pb = {k: torch.tensor(B[k], dtype=torch.float64, device="cuda") for k in qa.BEAM_ARRAYS}
    work = torch.empty(qa.beam_deposit_workspace_bytes(55000, nxi, s.M - 1), dtype=torch.uint8, device="cuda")
    bsrc = torch.zeros((nxi, s.M, nc), dtype=torch.float64, device="cuda")
    qa.beam_deposit_beams(g, qa.BeamGridC(xi_min, dxi, nxi, 0, nxi, 0), 1 if m1 else 0, [-1.0, 2.0], [pa, pb], bsrc, work)
    torch.cuda.synchronize()
    got = bsrc.cpu().numpy()
    assert _same(got, ref), _report("bsrc", got, ref)
    # rows 40..59 only, into a window array: the same rows
    win = torch.zeros((20, s.M, nc), dtype=torch.float64, device="cuda")
    qa.beam_deposit_beams(g, qa.BeamGridC(xi_min, dxi, nxi, 40, 20, 0), 1 if m1 else 0, [-1.0, 2.0], [pa, pb], win, work)
    torch.cuda.synchronize()
    got = win.cpu().numpy()
    assert _same(got, ref[40:60]), _report("window", got, ref[40:60])


@pytest.mark.parametrize("m1", [False, True])
@pytest.mark.parametrize("pusher", [0, 1, 2, 3])
@pytest.mark.parametrize("move", [False, True])
def test_beam_advance_bit_exact(torch_dev, qa, oracle, m1, pusher, move):
    torch = torch_dev
    rng = np.random.default_rng(100 + 8 * m1 + 2 * pusher + move)
    nodes = oracle.make_uniform(3.0, 0.02)
    s = oracle.Solver(nodes)
    nxi, dxi, xi_min = 200, 0.01, 0.0
    n = 20000
    parts = _beam_parts(rng, n, 2.9, 0.1, nxi * dxi - 0.1, m1)
    nc = 18 if m1 else 6
    fld = rng.normal(0, 1.0, (nxi, s.M, nc))
    ref = oracle.beam_advance(s, xi_min, dxi, nxi, False, m1, -1.0, pusher, parts, fld, 10.0, 10.0, move)
    r, ih, iV = _grid_dev(torch, s)
    g = qa.radial_grid(r, ih, iV)
    box = qa.BeamGridC(xi_min, dxi, nxi, 0, nxi, 0)
    p = {k: torch.tensor(parts[k], dtype=torch.float64, device="cuda") for k in qa.BEAM_ARRAYS}
    f = torch.tensor(fld, dtype=torch.float64, device="cuda")
    qa.beam_advance(g, box, 1 if m1 else 0, -1.0, pusher, p, f, 10.0, 10.0, move)
    torch.cuda.synchronize()
    for k in qa.BEAM_ARRAYS:
        got = p[k].cpu().numpy()
        assert _same(got, ref[k]), _report(k, got, ref[k])


# ------------------------------------------------------------------------------------------ whole step
def _compare_sims(a, b, what=("fld", "diag", "bsrc")):
    for w in what:
        ga, gb = getattr(a, w), getattr(b, w)
        assert _same(ga, gb), _report(w, ga, gb)


def test_sweep_blowout_m0_bit_exact(qa, oracle, decks):
    """The reference's blowout validation deck (m = 0, 3200 particles per slice, 1001 slices, near-singular E_z)."""
    ov = {"xi.min": -10}
    g = qa.Simulation(f"{decks}/blowout.in", ov)
    o = oracle.Sim(f"{decks}/blowout.in", ov)
    g.step()
    o.step()
    _compare_sims(g, o)


def test_sweep_stretched_m0_bit_exact(qa, oracle, decks):
    ov = {"xi.min": -10, "xi.step": 0.005, "grid.type": "regions", "grid.regions": "0.5:0.0025 2.5:0.01 8:0.05",
          "grid.max_ratio": 1.05}
    g = qa.Simulation(f"{decks}/blowout.in", ov)
    o = oracle.Sim(f"{decks}/blowout.in", ov)
    g.step()
    o.step()
    _compare_sims(g, o)


def test_sweep_m1_mobile_bit_exact(qa, oracle, decks):
    """m = 1 with mobile electrons (8 azimuths per ring), an offset driver and Picard coupling."""
    ov = {"modes": 1, "xi.min": -6, "grid.dr": 0.04, "driver.n0": 0.5, "driver.x0": 0.05, "driver.nparticles": 40000,
          "driver.sigma_r": 0.5, "electrons.ppc": 1}
    g = qa.Simulation(f"{decks}/linear.in", ov)
    o = oracle.Sim(f"{decks}/linear.in", ov)
    g.step()
    o.step()
    _compare_sims(g, o)


def test_steps_gain_bit_exact(qa, oracle, decks):
    """Driver + witness, beam pushes between sweeps: fields and witness phase space after 3 steps."""
    ov = {"time.steps": 3, "driver.nparticles": 50000, "xi.min": -8}
    g = qa.Simulation(f"{decks}/gain.in", ov)
    o = oracle.Sim(f"{decks}/gain.in", ov)
    for _ in range(4):
        g.step()
        o.step()
    _compare_sims(g, o)
    for k in qa.BEAM_ARRAYS:
        a, b = g.get(f"beam1.{k}"), o.get(f"beam1.{k}")
        assert _same(a, b), _report(f"witness {k}", a, b)
    # the beam log values: the product reduces the sums on the device, in a fixed tree order the oracle restates
    da, db = g.beam_diag(1), o.beam_diag(1, order="tree")
    for k in da:
        assert da[k] == db[k], (k, da[k], db[k])
    # the reference's index-order sums (host copy) bit for bit
    assert np.array_equal(g.beam_sums(1)[0], o.beam_sums(1)[0])


def test_betatron_m1_steps_bit_exact(qa, oracle, decks):
    """Immobile ion channel, offset witness, m = 1, 10 Vay steps."""
    ov = {"modes": 1, "witness.x0": 0.05, "witness.nsym": 4, "time.steps": 10}
    g = qa.Simulation(f"{decks}/betatron.in", ov)
    o = oracle.Sim(f"{decks}/betatron.in", ov)
    for _ in range(11):
        g.step()
        o.step()
    _compare_sims(g, o)
    for k in qa.BEAM_ARRAYS:
        a, b = g.get(f"beam0.{k}"), o.get(f"beam0.{k}")
        assert _same(a, b), _report(f"witness {k}", a, b)


@pytest.mark.parametrize("deck,ov,steps", [
    ("gain.in", {"time.steps": 2, "driver.nparticles": 20000, "xi.min": -8}, 2),
    ("betatron.in", {"modes": 1, "witness.x0": 0.05, "witness.nsym": 4, "time.steps": 2}, 2),
])
def test_fields_file_equals_oracle(qa, oracle, decks, tmp_path, deck, ov, steps):
    """fields_*.bin v3 written from HBM one component at a time: the same bytes as the oracle's writer (m = 0
    and m = 1, whose _c / _s and vector parts are formed on the device)."""
    g = qa.Simulation(f"{decks}/{deck}", ov)
    o = oracle.Sim(f"{decks}/{deck}", ov)
    for _ in range(steps):
        g.step()
        o.step()
    pg, po = tmp_path / "gpu.bin", tmp_path / "oracle.bin"
    g.write_fields(str(pg))
    o.write_fields(str(po))
    a, b = pg.read_bytes(), po.read_bytes()
    assert len(a) == len(b) and a == b, (len(a), len(b))


@pytest.mark.parametrize("deck,ov", [
    ("gain.in", {"time.steps": 2, "driver.nparticles": 20000, "xi.min": -8}),
    ("betatron.in", {"modes": 1, "witness.x0": 0.05, "witness.nsym": 4, "time.steps": 2}),
])
def test_axis_file_equals_oracle(qa, oracle, decks, tmp_path, deck, ov):
    """axis_*.txt, written with every field output from the first two nodes of each slice: the oracle's text
    byte for byte, after the time line of the header (the run writes it before the step's beam push, at t of
    the step; the oracle's file is made after the whole step)."""
    g = qa.Simulation(f"{decks}/{deck}", {**ov, "output.dir": str(tmp_path), "output.every": 1})
    o = oracle.Sim(f"{decks}/{deck}", ov)
    dt = None
    for n in range(3):
        g.step()
        o.step()
        ref = tmp_path / f"oracle_{n}.txt"
        o.write_axis(str(ref))
        a = (tmp_path / f"axis_{n:06d}.txt").read_bytes()
        b = ref.read_bytes()
        assert len(a) == len(b) and a[128:] == b[128:], (n, len(a), len(b))
        line = a[:128].decode()
        assert line.startswith("# t = ") and len(line) == 128 and line.endswith("\n"), line
        t = float(line[6:])
        if n == 1:
            dt = t
        if n == 2:
            assert t == 2 * dt, (t, dt)


def test_beam_log_equals_oracle(qa, oracle, decks, tmp_path):
    """beams.txt (Simulation::write_beam_output, one line per beam and step, after the sweep and before the
    push): the header, and every value equal to the oracle's beam log of the same beam state at 10 digits (the
    sums reduced on the device, in the tree order the oracle restates)."""
    ov = {"time.steps": 3, "driver.nparticles": 20000, "xi.min": -8}
    g = qa.Simulation(f"{decks}/gain.in", {**ov, "output.dir": str(tmp_path)})
    o = oracle.Sim(f"{decks}/gain.in", ov)
    want = []
    for n in range(3):
        # the oracle's beams as the GPU run logs them: after step n - 1's push, at the time of step n
        for b in range(o.nbeams):
            d = o.beam_diag(b, order="tree")
            want.append((n, b, d))
        g.step()
        o.step()
    lines = (tmp_path / "beams.txt").read_text().splitlines()
    assert lines[0] == ("# beam step t alive npart charge gamma_mean gamma_rms xi_mean xi_rms r_rms emit_nx"
                        " x_mean y_mean emit_ny sigma_x sigma_y")
    rows = [l.split() for l in lines[1:]]
    assert len(rows) == len(want)
    keys = ["t", "alive", "npart", "charge", "gamma_mean", "gamma_rms", "xi_mean", "xi_rms", "r_rms", "emit_nx",
            "x_mean", "y_mean", "emit_ny", "sigma_x", "sigma_y"]
    for r, (n, b, d) in zip(rows, want):
        assert int(r[1]) == n, r
        for k, v in zip(keys, r[2:]):
            ref = int(d[k]) if k == "alive" else d[k]
            assert v == (str(ref) if k == "alive" else f"{ref:.10g}"), (n, b, k, v, ref)


def test_torch_keeps_the_device_after_the_library(qa):
    """In a fresh process the package loads its library and steps a run before torch touches the GPU;
    torch must still see the device and compute (one HIP runtime in the process)."""
    import subprocess
    import sys
    import os
    code = ("import quarz_amd, sys\n"
            "from quarz_amd import configs\n"
            "s = quarz_amd.Simulation(configs.c4_deck(nd=2000, nw=1000), {'xi.min': -11.99})\n"
            "s.step()\n"
            "import torch\n"
            "ok = torch.cuda.is_available() and float(torch.ones(3, device='cuda').sum()) == 3.0\n"
            "print('torch ok' if ok else 'torch lost the device', flush=True)\n"
            "sys.exit(0 if ok else 1)\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]


def test_c1_workload_bit_exact(qa, oracle):
    """The c1 workload at full size and length (m = 0, 256 uniform cells, 1601 slices, 1e5-particle driver, its
    10 time steps: 11 sweeps and 10 pushes): fields, diagnostics, beam sources, the driver's phase space and its
    beam-log moments after the 10 steps equal the oracle's bit for bit."""
    from quarz_amd import configs
    deck = configs.deck("c1")
    g = qa.Simulation(deck)
    o = oracle.Sim(deck)
    assert g.nsteps == 10
    for _ in range(11):
        g.step()
        o.step()
    _compare_sims(g, o)
    for k in qa.BEAM_ARRAYS:
        a, b = g.get(f"beam0.{k}"), o.get(f"beam0.{k}")
        assert _same(a, b), _report(f"driver {k}", a, b)
    da, db = g.beam_diag(0), o.beam_diag(0, order="tree")
    for k in da:
        assert da[k] == db[k], (k, da[k], db[k])
    # the reference's index-order moments after the 10 steps, bit for bit; the device order differs from them
    # only in the last bits (a 1e-9 relative gate on beam moments after 10 steps is the parity bar)
    sg, so = g.beam_sums(0)[0], o.beam_sums(0)[0]
    assert np.array_equal(sg, so)
    di = o.beam_diag(0)
    for k in ("gamma_mean", "gamma_rms", "r_rms", "emit_nx", "xi_mean", "xi_rms"):
        assert abs(da[k] - di[k]) <= 1e-12 * abs(di[k]), (k, da[k], di[k])


def test_thermal_plasma_bit_exact(qa, oracle, decks):
    """A warm plasma (electrons.uth > 0: momenta drawn from std::normal_distribution at the first load, the same
    every step) keeps the host load and its device copies; a cold one is loaded by the sweep kernel itself (every
    other test).  Three steps bit for bit."""
    ov = {"time.steps": 2, "driver.nparticles": 20000, "xi.min": -6, "electrons.uth": 0.01}
    g = qa.Simulation(f"{decks}/gain.in", ov)
    o = oracle.Sim(f"{decks}/gain.in", ov)
    for _ in range(3):
        g.step()
        o.step()
    _compare_sims(g, o)
    for k in ("x", "y", "px", "py", "dl", "w"):
        a, b = g.get(f"sp0.{k}"), o.get(f"sp0.{k}")
        assert _same(a, b), _report(f"electrons {k}", a, b)


@pytest.mark.parametrize("m1", [False, True])
def test_two_mobile_species_bit_exact(qa, oracle, decks, m1):
    """Electrons and mobile hydrogen ions (two species in one sweep: the positive species' deposit into n_i, the
    split of the particle range between species, the second species' push).  m = 0: the shape of the
    reference's ion-motion validation (examples/pwfa_ion_motion.in, stretched 512 cells, plasma to r = 7, ppc 4
    each, about 3.9e3 particles per slice) with smaller beams, three steps.  m = 1: an offset driver, 8 azimuths
    per ring for both species, three steps."""
    if not m1:
        deck = f"{decks}/pwfa_ion_motion.in"
        ov = {"xi.min": -7.3, "time.steps": 2, "driver.nparticles": 40000, "witness.nparticles": 20000}
    else:
        deck = f"{decks}/linear.in"
        ov = {"modes": 1, "xi.min": -5, "grid.dr": 0.04, "driver.n0": 0.5, "driver.x0": 0.05,
              "driver.nparticles": 20000, "driver.sigma_r": 0.5, "driver.rigid": 0, "time.steps": 2, "time.dt": 5,
              "plasma.species": "electrons ions", "electrons.ppc": 1, "ions.ppc": 1, "ions.charge": 1,
              "ions.mass": 1836.15, "ions.mobile": 1}
    g = qa.Simulation(deck, ov)
    o = oracle.Sim(deck, ov)
    assert o.nspecies == 2
    for _ in range(3):
        g.step()
        o.step()
    _compare_sims(g, o)
    for bi in range(o.nbeams):
        for k in qa.BEAM_ARRAYS:
            a, b = g.get(f"beam{bi}.{k}"), o.get(f"beam{bi}.{k}")
            assert _same(a, b), _report(f"beam{bi} {k}", a, b)
    for sp in range(2):
        for k in ("x", "y", "px", "py", "dl", "w"):
            a, b = g.get(f"sp{sp}.{k}"), o.get(f"sp{sp}.{k}")
            assert _same(a, b), _report(f"species {sp} {k}", a, b)
    assert g.lost() == o.lost()


def test_c2_workload_pipelined_bit_exact(qa, oracle, tmp_path):
    """The c2 workload at full size (m = 0 + 1, 512 stretched cells, 4001 slices, driver 1e6 + witness 5e5)
    through the bench's path: the start-up step, then two steps in flight over 128 slice ranges.  Fields,
    diagnostics, beam sources and both beams equal the oracle's three serial steps bit for bit."""
    from quarz_amd import configs
    deck = configs.deck("c2", workdir=str(tmp_path))
    g = qa.Simulation(deck, {"time.steps": 4})
    o = oracle.Sim(deck, {"time.steps": 4})
    g.step()
    g.run_pipelined(2, chunks=128, depth=2)
    for _ in range(3):
        o.step()
    _compare_sims(g, o)
    for bi in range(2):
        for k in qa.BEAM_ARRAYS:
            a, b = g.get(f"beam{bi}.{k}"), o.get(f"beam{bi}.{k}")
            assert _same(a, b), _report(f"beam{bi} {k}", a, b)


def test_c5_runs_batched_bit_exact(qa, oracle, tmp_path):
    """Two runs of the c5 scan (their own grids, slice counts, beams), stepped as a Batch (eager start-up
    step, eager step that captures, one graph replay): each run equals the oracle's serial run bit for bit."""
    from quarz_amd import configs
    runs = configs.c5_runs(nruns=2, steps=3, workdir=str(tmp_path))
    sims = [qa.Simulation(deck) for deck, _ in runs]
    batch = qa.Batch(sims, graph=True)
    for _ in range(3):
        batch.step()
    replays = batch.replays
    batch.close()
    assert replays >= 1
    for i, ((deck, info), g) in enumerate(zip(runs, sims)):
        o = oracle.Sim(deck)
        for _ in range(3):
            o.step()
        for w in ("fld", "diag", "bsrc"):
            a, b = getattr(g, w), getattr(o, w)
            assert _same(a, b), _report(f"run {i} ({info}) {w}", a, b)
        for bi in range(2):
            for k in qa.BEAM_ARRAYS:
                a, b = g.get(f"beam{bi}.{k}"), o.get(f"beam{bi}.{k}")
                assert _same(a, b), _report(f"run {i} beam{bi} {k}", a, b)


def test_c5_bench_batch_extremes_equal_plain(qa, tmp_path):
    """The bench's c5 batch: all 256 runs balanced by the batch autotuner (workgroups per run), then stepped as a
    Batch with graph replays (the eager start-up step, then four replays, the first of which captures the graph)
    for the workload's 5 steps.  The runs at the extremes (fewest and most cells, fewest and most slices) and the
    two shortest and two longest of the runs the plan gave several workgroups equal the same decks stepped one
    run at a time on one workgroup, bit for bit."""
    from quarz_amd import configs
    runs = configs.c5_runs(256, steps=7, workdir=str(tmp_path))
    sims = [qa.Simulation(deck) for deck, _ in runs]
    batch = qa.Batch(sims, graph=True)
    try:
        plan = batch.balance()
        print("plan: est", plan["est_ms"], "ms, one workgroup per run", plan["est_one_ms"], "ms; groups",
              {g: plan["groups"].count(g) for g in sorted(set(plan["groups"]))})
        for _ in range(5):
            batch.step()
        replays = batch.replays
    finally:
        batch.close()
    assert replays == 4
    info = [i for _, i in runs]
    pick = sorted({min(range(256), key=lambda i: (info[i]["N"], info[i]["K"])),
                   max(range(256), key=lambda i: (info[i]["N"], info[i]["K"])),
                   min(range(256), key=lambda i: (info[i]["K"], info[i]["N"])),
                   max(range(256), key=lambda i: (info[i]["K"], info[i]["N"]))})
    multi = sorted((i for i in range(256) if plan["groups"][i] > 1), key=lambda i: (info[i]["K"], info[i]["N"]))
    pick = sorted(set(pick) | set(multi[:2] + multi[-2:]))
    for i in pick:
        p = qa.Simulation(runs[i][0])
        for _ in range(5):
            p.step()
        for w in ("fld", "diag", "bsrc"):
            a, b = getattr(sims[i], w), getattr(p, w)
            assert _same(a, b), _report(f"run {i} ({info[i]}) {w}", a, b)
        for bi in range(2):
            for k in qa.BEAM_ARRAYS:
                a, b = sims[i].get(f"beam{bi}.{k}"), p.get(f"beam{bi}.{k}")
                assert _same(a, b), _report(f"run {i} beam{bi} {k}", a, b)
        assert sims[i].lost() == p.lost()
        del p


# ------------------------------------------------------------------------------------------ batched runs
def _batched_vs_plain(qa, deck, overrides, steps):
    plain = [qa.Simulation(deck, ov) for ov in overrides]
    batch = [qa.Simulation(deck, ov) for ov in overrides]
    for _ in range(steps):
        for s in plain:
            s.step()
        qa.step_batched(batch)
    for i, (a, b) in enumerate(zip(batch, plain)):
        for w in ("fld", "diag", "bsrc"):
            ga, gb = getattr(a, w), getattr(b, w)
            assert _same(ga, gb), _report(f"run {i} {w}", ga, gb)
        for bi in range(a.nbeams):
            for k in qa.BEAM_ARRAYS:
                ga, gb = a.get(f"beam{bi}.{k}"), b.get(f"beam{bi}.{k}")
                assert _same(ga, gb), _report(f"run {i} beam{bi}.{k}", ga, gb)
        assert a.lost() == b.lost()


def test_batched_m0_equals_plain(qa, decks):
    """Runs of different grids, slice counts and particle counts in one launch: each bit-identical to its
    own plain launch (the batched form adds no cross-run interaction)."""
    ov = [{"time.steps": 3, "driver.nparticles": 20000, "xi.min": -8},
          {"time.steps": 3, "driver.nparticles": 30000, "xi.min": -6, "grid.type": "uniform", "grid.dr": 0.02,
           "electrons.ppc": 2},
          {"time.steps": 3, "driver.nparticles": 10000, "xi.min": -9, "grid.type": "uniform", "grid.dr": 0.08,
           "electrons.ppc": 8}]
    _batched_vs_plain(qa, f"{decks}/gain.in", ov, 3)


def test_batched_m1_equals_plain(qa, decks):
    ov = [{"modes": 1, "xi.min": -5, "grid.dr": 0.04, "driver.n0": 0.5, "driver.x0": 0.05,
           "driver.nparticles": 20000, "driver.sigma_r": 0.5, "electrons.ppc": 1, "time.steps": 2, "time.dt": 5},
          {"modes": 1, "xi.min": -4, "grid.dr": 0.05, "driver.n0": 0.3, "driver.x0": 0.1,
           "driver.nparticles": 10000, "driver.sigma_r": 0.4, "electrons.ppc": 1, "time.steps": 2, "time.dt": 5}]
    _batched_vs_plain(qa, f"{decks}/linear.in", ov, 2)


def _graph_vs_plain(qa, deck, overrides, steps, replays):
    """A Batch with graph replays against the same runs stepped plainly, compared bit for bit after every
    step: the eager start-up step, the step that captures the graph, each replay (with the host work of the
    step and the comparison's downloads between replays) and the eager last step."""
    plain = [qa.Simulation(deck, ov) for ov in overrides]
    runs = [qa.Simulation(deck, ov) for ov in overrides]
    batch = qa.Batch(runs, graph=True)
    try:
        for st in range(steps):
            for s in plain:
                s.step()
            batch.step()
            for i, (a, b) in enumerate(zip(runs, plain)):
                for w in ("fld", "diag", "bsrc"):
                    ga, gb = getattr(a, w), getattr(b, w)
                    assert _same(ga, gb), _report(f"step {st} run {i} {w}", ga, gb)
                for bi in range(a.nbeams):
                    for k in qa.BEAM_ARRAYS:
                        ga, gb = a.get(f"beam{bi}.{k}"), b.get(f"beam{bi}.{k}")
                        assert _same(ga, gb), _report(f"step {st} run {i} beam{bi}.{k}", ga, gb)
                assert a.lost() == b.lost()
                assert a.t == b.t
        assert batch.replays == replays
    finally:
        batch.close()


def test_graph_batch_m0_equals_plain(qa, decks):
    """time.steps = 5: step 0 starts the leapfrog (eager), steps 1-4 are graph replays (the first of them
    captures), step 5 is the last sweep without a push (eager)."""
    ov = [{"time.steps": 5, "driver.nparticles": 20000, "xi.min": -8},
          {"time.steps": 5, "driver.nparticles": 30000, "xi.min": -6, "grid.type": "uniform", "grid.dr": 0.02,
           "electrons.ppc": 2},
          {"time.steps": 5, "driver.nparticles": 10000, "xi.min": -9, "grid.type": "uniform", "grid.dr": 0.08,
           "electrons.ppc": 8}]
    _graph_vs_plain(qa, f"{decks}/gain.in", ov, 6, 4)


def test_graph_batch_m1_equals_plain(qa, decks):
    ov = [{"modes": 1, "xi.min": -5, "grid.dr": 0.04, "driver.n0": 0.5, "driver.x0": 0.05,
           "driver.nparticles": 20000, "driver.sigma_r": 0.5, "electrons.ppc": 1, "time.steps": 4, "time.dt": 5},
          {"modes": 1, "xi.min": -4, "grid.dr": 0.05, "driver.n0": 0.3, "driver.x0": 0.1,
           "driver.nparticles": 10000, "driver.sigma_r": 0.4, "electrons.ppc": 1, "time.steps": 4, "time.dt": 5}]
    _graph_vs_plain(qa, f"{decks}/linear.in", ov, 5, 3)


# ------------------------------------------------------------------------------------------ pipelined steps
@pytest.mark.parametrize("deck,ov,steps,chunks,depth", [
    ("gain.in", {"time.steps": 4, "driver.nparticles": 20000, "xi.min": -8}, 4, 5, 3),
    ("gain.in", {"time.steps": 6, "driver.nparticles": 20000, "xi.min": -8, "beams.xi_shape": "ngp"}, 3, 7, 2),
    ("gain.in", {"time.steps": 3, "driver.nparticles": 10000, "xi.min": -6, "grid.type": "uniform",
                 "grid.dr": 0.05}, 3, 1, 1),
    ("linear.in", {"modes": 1, "xi.min": -5, "grid.dr": 0.04, "driver.n0": 0.5, "driver.x0": 0.05,
                   "driver.nparticles": 20000, "driver.sigma_r": 0.5, "electrons.ppc": 1, "time.steps": 4,
                   "time.dt": 5}, 4, 4, 4),
    # a slow witness (gamma 3) slips about 100 slices per step, across several ranges: a range's deposit must
    # not see the particles its own step has already pushed from earlier ranges
    ("gain.in", {"time.steps": 4, "driver.nparticles": 20000, "xi.min": -8, "witness.gamma": 3,
                 "witness.sigma_xi": 1.0, "witness.xi0": -4, "witness.n0": 0.001}, 3, 32, 3),
])
def test_pipelined_steps_equal_plain(qa, decks, deck, ov, steps, chunks, depth):
    """Steps overlapped in time (step n sweeps a slice range once step n-1 has pushed the beam particles that
    reach it) against the same steps taken one at a time: fields, diagnostics, the last beam sources, every
    beam particle, the lost count and the time, bit for bit.  The cases end on the last sweep (no push) or
    before it, with one to four steps in flight and one to seven ranges per step."""
    a = qa.Simulation(f"{decks}/{deck}", ov)
    b = qa.Simulation(f"{decks}/{deck}", ov)
    a.step()
    b.step()   # the leapfrog start-up step
    psl = pushed = 0.0
    for i in range(steps):
        a.step()   # step 1 + i; the last one (1 + i == nsteps) has no push
        st = a.stats()
        psl += st["particle_slices"]
        pushed += st["beam_pushed"] if 1 + i < a.nsteps else 0.0
    bpsl, bpushed = b.run_pipelined(steps, chunks=chunks, depth=depth)
    # the work counts of the overlapped steps are the serial ones (live particles before each deposit and push)
    assert (bpsl, bpushed) == (psl, pushed)
    assert b.stats()["beam_deposited"] == a.stats()["beam_deposited"]
    for w in ("fld", "diag", "bsrc"):
        ga, gb = getattr(b, w), getattr(a, w)
        assert _same(ga, gb), _report(f"pipelined {w}", ga, gb)
    for bi in range(a.nbeams):
        for k in qa.BEAM_ARRAYS:
            ga, gb = b.get(f"beam{bi}.{k}"), a.get(f"beam{bi}.{k}")
            assert _same(ga, gb), _report(f"pipelined beam{bi}.{k}", ga, gb)
    assert b.lost() == a.lost()
    assert b.t == a.t


def test_pipelined_with_many_hardware_queues(qa):
    """The bench's setting: 24 hardware queues, so that the steps in flight really overlap (in this process
    HIP has started with its default of 4).  A child process compares pipelined and plain steps for a witness
    that slips several ranges per step and for the c2 workload (four steps in flight)."""
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    env = dict(os.environ, GPU_MAX_HW_QUEUES="24")
    r = subprocess.run([sys.executable, os.path.join(here, "pipelined_check.py")], capture_output=True, text=True,
                       timeout=280, env=env)
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]


def test_pipelined_betatron_m1_equals_oracle(qa, oracle, decks):
    """Pipelined steps against the serial oracle directly: the betatron deck (m = 1, offset witness), the
    start-up step, then 10 steps with 5 in flight over 6 slice ranges, ending on the last sweep."""
    ov = {"modes": 1, "witness.x0": 0.05, "witness.nsym": 4, "time.steps": 10}
    g = qa.Simulation(f"{decks}/betatron.in", ov)
    o = oracle.Sim(f"{decks}/betatron.in", ov)
    g.step()
    for _ in range(11):
        o.step()
    g.run_pipelined(10, chunks=6, depth=5)
    _compare_sims(g, o)
    for k in qa.BEAM_ARRAYS:
        a, b = g.get(f"beam0.{k}"), o.get(f"beam0.{k}")
        assert _same(a, b), _report(f"witness {k}", a, b)


# ------------------------------------------------------------------------------------------ launch shapes
def _state(qa, sim):
    out = {w: getattr(sim, w) for w in ("fld", "diag", "bsrc")}
    for bi in range(sim.nbeams):
        for k in qa.BEAM_ARRAYS:
            out[f"beam{bi}.{k}"] = sim.get(f"beam{bi}.{k}")
    return out


def test_launch_shapes_and_autotune_keep_results(qa, decks):
    """Beam kernels at 64 and 1024 threads per block give the results of the default 256 bit for bit; the
    autotuner leaves the run's state untouched and picks one of its candidates."""
    ov = {"time.steps": 3, "driver.nparticles": 30000, "xi.min": -8}
    ref = qa.Simulation(f"{decks}/gain.in", ov)
    for _ in range(2):
        ref.step()
    want = _state(qa, ref)
    try:
        for threads in (64, 1024):
            for k in qa.KERNELS:
                qa.set_block(k, threads)
            s = qa.Simulation(f"{decks}/gain.in", ov)
            for _ in range(2):
                s.step()
            got = _state(qa, s)
            for name in want:
                assert _same(got[name], want[name]), _report(f"{threads} threads {name}", got[name], want[name])
        for k in qa.KERNELS:
            qa.set_block(k, 256)
        s = qa.Simulation(f"{decks}/gain.in", ov)
        s.step()
        before = _state(qa, s)
        chosen = s.autotune(candidates=(64, 256), reps=1)
        after = _state(qa, s)
        for name in before:
            assert _same(after[name], before[name]), _report(f"autotune changed {name}", after[name], before[name])
        assert set(chosen) == set(qa.KERNELS) and all(v in (64, 256) for v in chosen.values())
        s.step()
        got = _state(qa, s)
        for name in want:
            assert _same(got[name], want[name]), _report(f"tuned {name}", got[name], want[name])
    finally:
        for k in qa.KERNELS:
            qa.set_block(k, 256)


# ------------------------------------------------------------------------------------------ beam sums
@pytest.mark.parametrize("np_", [0, 1, 4095, 4097, 300001])
def test_beam_sums_device_matches_host(torch_dev, qa, np_):
    """On-device Beam::sums against the index-ordered host sums (bit-identical to the reference's loop):
    the live count exactly, every weighted sum to 1e-12 of the sum of its absolute terms."""
    torch = torch_dev
    rng = np.random.default_rng(np_ + 11)
    host = {k: rng.normal(0, 1, np_) for k in qa.BEAM_ARRAYS}
    host["pz"] = 2000 + 20 * rng.normal(0, 1, np_)
    host["w"] = rng.uniform(0.5, 1.5, np_) * (rng.uniform(0, 1, np_) > 0.1)
    p = {k: torch.tensor(v, dtype=torch.float64, device="cuda") for k, v in host.items()}
    gref, xiref = 2000.0, -2.0
    ref = qa.beam_sums(p, gref, xiref)
    out = torch.zeros(17, dtype=torch.float64, device="cuda")
    work = torch.zeros(max(1, qa.beam_sums_workspace_bytes(np_) // 8), dtype=torch.float64, device="cuda")
    qa.beam_sums_device(p, gref, xiref, out, work)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert got[16] == ref[16] == np.count_nonzero(host["w"])
    # absolute scale of every sum: the same sums over |terms|
    w = host["w"]
    g = np.sqrt(1 + host["px"] * host["px"] + host["py"] * host["py"] + host["pz"] * host["pz"])
    G, XI = g - gref, host["xi"] - xiref
    X, Y, PX, PY = host["x"], host["y"], host["px"], host["py"]
    terms = [np.ones_like(w), G, G * G, XI, XI * XI, X * X + Y * Y, X, Y, X * X, Y * Y, PX, PY, PX * PX, PY * PY,
             X * PX, Y * PY]
    for c, v in enumerate(terms):
        scale = np.abs(w * v).sum()
        assert abs(got[c] - ref[c]) <= 1e-12 * scale + 1e-300, (c, got[c], ref[c], scale)
    # and exactly the fixed tree order: blocks of 4096 particles, lane t of 256 accumulating t, t + 256, ... in
    # index order (dead particles skipped), lanes halved pairwise, block partials added in block order
    live = w != 0
    vals = np.stack([live.astype(np.float64), w] + [w * v for v in terms[1:]])   # 17 x np: count, w, w v
    tree = np.zeros(17)
    for lo in range(0, np_, 4096):
        acc = np.zeros((17, 256))
        for i0 in range(lo, min(np_, lo + 4096), 256):
            n = min(256, np_ - i0)
            blk = vals[:, i0:i0 + n]
            acc[:, :n] = np.where(live[i0:i0 + n], acc[:, :n] + blk, acc[:, :n])
        h = 128
        while h > 0:
            acc[:, :h] = acc[:, :h] + acc[:, h:2 * h]
            h //= 2
        tree = tree + acc[:, 0]
    order = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 0]   # device layout: w, w v..., count
    assert np.array_equal(got, tree[order]), (got, tree[order])
