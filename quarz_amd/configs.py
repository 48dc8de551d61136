"""The benchmark workloads as input decks (the reference's deck format, src/Config.cpp).

c1  m = 0 blowout, uniform 256 cells to R = 8, ppc 4, xi in [-16, 0] step 0.01 (1601 slices), 1e5-particle
    Gaussian driver, Vay pusher, dt 20, 10 steps.
c2  m = 0 + 1, stretched 512 cells to R = 10 (5e-4 at the axis, 0.05 at the wall), 8 azimuths per ring,
    xi in [-20, 0] step 0.005 (4001 slices), offset driver (1e6 particles) and witness (5e5), Picard 3,
    dt 10, 20 steps.
Driver parameters a workload leaves open (gamma, xi0, energy spread, emittance) are the c1 values.
"""
import os
import tempfile


def stretched_nodes(ncells, R, h0, hmax):
    """Radial nodes of `ncells` cells growing geometrically from h0 at the axis up to hmax, then uniform,
    scaled to end exactly at R.  The growth ratio q is found by bisection."""
    def total(q):
        s, h = 0.0, h0
        for _ in range(ncells):
            s += min(h, hmax)
            h *= q
        return s
    lo, hi = 1.0, 1.5
    if total(hi) < R:
        raise ValueError("cannot reach R with these cells")
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if total(mid) < R:
            lo = mid
        else:
            hi = mid
    q = 0.5 * (lo + hi)
    cells, h = [], h0
    for _ in range(ncells):
        cells.append(min(h, hmax))
        h *= q
    scale = R / sum(cells)
    nodes = [0.0]
    for c in cells:
        nodes.append(nodes[-1] + c * scale)
    nodes[-1] = R
    return nodes, q


_C1 = """
xi.convention = z-t
modes = 0
grid.type = uniform
grid.rmax = 8
grid.dr = 0.03125
xi.min = -16
xi.max = 0
xi.step = 0.01
time.dt = 20
time.steps = 10
pusher.beam = vay
pusher.ab_order = 3
plasma.species = electrons
electrons.ppc = 4
beams = driver
driver.n0 = 2
driver.sigma_r = 0.3
driver.sigma_xi = 1.0
driver.xi0 = -2
driver.gamma = 2000
driver.espread = 0.01
driver.emit_n = 0.002
driver.cut = 3
driver.nparticles = 100000
driver.seed = 12345
"""

_C2 = """
xi.convention = z-t
modes = 1
solver.picard = 3
grid.type = file
grid.file = {grid_file}
xi.min = -20
xi.max = 0
xi.step = 0.005
time.dt = 10
time.steps = 20
pusher.beam = vay
pusher.ab_order = 3
plasma.species = electrons
electrons.ppc = 1
electrons.ntheta = 8
beams = driver witness
driver.n0 = 4
driver.sigma_r = 0.2
driver.sigma_xi = 0.7
driver.x0 = 0.1
driver.xi0 = -2
driver.gamma = 2000
driver.espread = 0.01
driver.emit_n = 0.002
driver.cut = 3
driver.nparticles = 1000000
witness.n0 = 10
witness.sigma_r = 0.02
witness.sigma_xi = 0.2
witness.xi0 = -8
witness.gamma = 2000
witness.emit_n = 0
witness.cut = 3
witness.nparticles = 500000
"""

DESCRIPTION = {
    "c1": "m0 blowout, uniform 256 cells R=8, ppc4, 1601 slices, driver 1e5, vay, dt20",
    "c2": "m0+1 stretched 512 cells (5e-4..0.05) R=10, 8 azimuths, 4001 slices, driver 1e6 + witness 5e5, "
          "picard 3, vay, dt10",
}


def deck(name, workdir=None):
    """Deck text of a workload; the c2 grid file is written into workdir (a temporary directory if None)."""
    if name == "c1":
        return _C1
    if name == "c2":
        workdir = workdir or tempfile.mkdtemp(prefix="quarz_amd_")
        nodes, _ = stretched_nodes(512, 10.0, 5e-4, 0.05)
        path = os.path.join(workdir, "grid_c2.txt")
        with open(path, "w") as f:
            f.write("# 512 cells, 5e-4 at the axis, 0.05 at the wall, R = 10\n")
            f.writelines(f"{x:.17g}\n" for x in nodes)
        return _C2.format(grid_file=path)
    raise KeyError(name)
