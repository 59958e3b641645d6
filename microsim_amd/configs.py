"""Input generators and legacy-VTK helpers for the 2-D grand-potential solver.

Everything here produces files in the formats the serial ``microsim_gp`` driver reads and writes:

* ``Input.in``  -- ``KEY = value;`` lines, parsed line by line with
  ``sscanf(line, "%100s = %100[^;];", key, value)``
  (``Grand_potential_Finite_difference_2D_serial/functions/reading_input_parameters.h:24-25``).
  Keys that allocate tables (``NUMPHASES``, ``NUMCOMPONENTS``) must precede the keys that fill them, and the
  ``BOUNDARY`` lines must precede the ``BOUNDARY_VALUE`` lines, because the halo proxy tables are only filled
  while ``BOUNDARY_VALUE`` is parsed (``functions/utility_functions.h:422-470, 544-650``).
* ``Filling.in`` -- ``FILLCYLINDER = {phase, x, y, z0, z1, r};`` etc. (``functions/fill_domain.h:28-147``).
* legacy VTK STRUCTURED_POINTS, big-endian doubles, fields in the order phases, ``Mu_<comp>``
  (``solverloop/file_writer.h:151-255``); the restart reader skips 8 header lines and 2 lines per field
  (``solverloop/file_writer.h:328-416``).

The benchmark workloads (binary tanh seed, ternary random seeds) are built by the ``config_*`` functions; the
random multi-seed layouts are explicit ``FILLCYLINDER`` lines from a seeded Poisson-disc generator, because the
reference's own random fillers are not reproducible.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# ----------------------------------------------------------------------------------------------------------------
# Input.in
# ----------------------------------------------------------------------------------------------------------------


def _fmt(v) -> str:
    if isinstance(v, bool):
        return "1" if v else "0"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return repr(float(v))
    return str(v)


def _tuple(vals: Sequence) -> str:
    return "{" + ", ".join(_fmt(v) for v in vals) + "}"


@dataclass
class GPInput:
    """A complete Input.in description for the 2-D GP solver (serial key set)."""

    mesh_x: int
    mesh_y: int
    nphases: int
    ncomp: int
    dx: float = 1.0
    dy: float = 1.0
    dt: float = 0.05
    ntimesteps: int = 100
    savet: int = 100
    nsmooth: int = 1
    starttime: int = 0
    restart: int = 0
    track_progress: int = 100
    writeformat: str = "BINARY"
    components: Sequence[str] = ()
    phases: Sequence[str] = ()
    epsilon: float = 4.0
    tau: float = 0.3
    gamma: Sequence[float] = ()          # upper triangle (0,1),(0,2),..,(1,2),..
    tau_ab: Sequence[float] = ()         # upper triangle
    dab: Sequence[float] = ()            # upper triangle
    gamma_abc: Sequence[float] = ()      # i<j<k triples
    fn_anisotropy: int = 1
    anisotropy_type: int = 4
    rotations: Dict[Tuple[int, int], Tuple[float, float, float]] = field(default_factory=dict)
    fn_w: int = 1
    diffusivity: Dict[int, Sequence[float]] = field(default_factory=dict)  # phase -> diagonal entries
    A: Dict[int, Sequence[float]] = field(default_factory=dict)            # phase -> (A11..Ann, A12, ...)
    ceq: Dict[Tuple[int, int], Sequence[float]] = field(default_factory=dict)
    cfill: Dict[Tuple[int, int], Sequence[float]] = field(default_factory=dict)
    slopes: Dict[Tuple[int, int], Sequence[float]] = field(default_factory=dict)
    T: float = 1.0
    Teq: float = 1.0
    Tfill: float = 1.0
    solve: str = "auto"                  # BINARY / TERNARY / DILUTE / LU / auto
    bc_phi: Tuple[int, int, int, int] = (1, 1, 1, 1)   # X+, X-, Y+, Y-  (1 Neumann, 3 periodic)
    bc_mu: Optional[Tuple[int, int, int, int]] = None
    writecomposition: int = 0
    shift: int = 0
    shiftj: int = 30
    # ISOTHERMAL = 0 with Tempgrady = {base_temp, DeltaT, Distance, gradient_OFFSET, velocity}; None: isothermal
    tempgrady: Optional[Tuple[float, float, float, float, float]] = None
    extra: List[Tuple[str, str]] = field(default_factory=list)

    def text(self) -> str:
        N, K = self.nphases, self.ncomp
        comps = list(self.components) or ["C%d" % i for i in range(K)]
        phases = list(self.phases) or (["S%d" % i for i in range(N - 1)] + ["L"])
        bc_mu = self.bc_mu if self.bc_mu is not None else self.bc_phi
        L: List[str] = []
        add = lambda k, v: L.append("%s = %s;" % (k, v))
        add("DIMENSION", 2)
        add("MESH_X", self.mesh_x)
        add("MESH_Y", self.mesh_y)
        add("MESH_Z", 1)
        add("DELTA_X", _fmt(self.dx))
        add("DELTA_Y", _fmt(self.dy))
        add("DELTA_Z", _fmt(self.dx))
        add("DELTA_t", _fmt(self.dt))
        add("NUMPHASES", N)
        add("NUMCOMPONENTS", K)
        add("NTIMESTEPS", self.ntimesteps)
        add("NSMOOTH", self.nsmooth)
        add("SAVET", self.savet)
        add("STARTTIME", self.starttime)
        add("RESTART", self.restart)
        add("COMPONENTS", _tuple(comps))
        add("PHASES", _tuple(phases))
        add("GAMMA", _tuple(self.gamma))
        for p in sorted(self.diffusivity):
            add("DIFFUSIVITY", _tuple([1, p] + list(self.diffusivity[p])))
        add("R", 1.0)
        add("V", 1.0)
        for side in ("phi", "mu", "c", "T"):
            b = bc_mu if side in ("mu", "c") else self.bc_phi
            add("BOUNDARY", _tuple([side] + list(b) + [0, 0]))
        for side in ("phi", "mu", "c", "T"):
            add("BOUNDARY_VALUE", _tuple([side, 0, 0, 0, 0, 0, 0]))
        if self.tempgrady is None:
            add("ISOTHERMAL", 1)
        else:
            add("ISOTHERMAL", 0)
            add("Tempgrady", _tuple(self.tempgrady))
        solve = self.solve
        if solve == "auto":
            solve = {2: "BINARY", 3: "TERNARY"}.get(K, "LU")
        if solve == "BINARY":
            add("BINARY", 1)
        elif solve == "TERNARY":
            add("TERNARY", 1)
        elif solve == "DILUTE":
            add("DILUTE", 1)
        add("T", _fmt(self.T))
        add("WRITEFORMAT", self.writeformat)
        add("TRACK_PROGRESS", self.track_progress)
        add("epsilon", _fmt(self.epsilon))
        add("tau", _fmt(self.tau))
        add("Tau", _tuple(self.tau_ab))
        add("Function_anisotropy", self.fn_anisotropy)
        add("Anisotropy_type", self.anisotropy_type)
        add("dab", _tuple(self.dab))
        if self.fn_anisotropy:
            for a in range(N):
                for b in range(a + 1, N):
                    th = self.rotations.get((a, b), (0.0, 0.0, 0.0))
                    add("Rotation_matrix", _tuple([a, b] + list(th)))
        add("Function_W", self.fn_w)
        add("Gamma_abc", _tuple(self.gamma_abc))
        add("Shift", self.shift)
        add("Shiftj", self.shiftj)
        add("Writecomposition", self.writecomposition)
        add("Noise_phasefield", 0)
        add("Amp_Noise_Phase", 0.0)
        add("Equilibrium_temperature", _fmt(self.Teq))
        add("Filling_temperature", _fmt(self.Tfill))
        add("Function_F", 1)
        for p in sorted(self.A):
            add("A", _tuple([p] + list(self.A[p])))
        for (a, b) in sorted(self.ceq):
            add("ceq", _tuple([a, b] + list(self.ceq[(a, b)])))
        for (a, b) in sorted(self.cfill):
            add("cfill", _tuple([a, b] + list(self.cfill[(a, b)])))
        for (a, b) in sorted(self.slopes):
            add("slopes", _tuple([a, b] + list(self.slopes[(a, b)])))
        for k, v in self.extra:
            add(k, v)
        return "\n".join(L) + "\n"


@dataclass
class Fill:
    kind: str          # "cylinder" | "cube" | "sphere"
    phase: int
    args: Tuple

    def line(self) -> str:
        if self.kind == "cylinder":
            x, y, r = self.args
            return "FILLCYLINDER = %s;" % _tuple([self.phase, int(x), int(y), 0, 0, _fmt(float(r))])
        if self.kind == "cube":
            x0, y0, x1, y1 = self.args
            return "FILLCUBE = %s;" % _tuple([self.phase, int(x0), int(y0), 0, int(x1), int(y1), 0])
        if self.kind == "sphere":
            x, y, r = self.args
            return "FILLSPHERE = %s;" % _tuple([self.phase, int(x), int(y), 0, _fmt(float(r))])
        raise ValueError(self.kind)


def filling_text(fills: Sequence[Fill]) -> str:
    return "\n".join(f.line() for f in fills) + "\n"


# ----------------------------------------------------------------------------------------------------------------
# Seeds and material tables
# ----------------------------------------------------------------------------------------------------------------


def poisson_disc_seeds(rng: np.random.Generator, mesh_x: int, mesh_y: int, n: int, rmin: float, rmax: float,
                       min_spacing: float, solid_phases: int, margin: int = 0, max_tries: int = 200000) -> List[Fill]:
    """Dart-throwing Poisson-disc sampler: ``n`` circular seeds with centre spacing >= ``min_spacing``.

    Each seed gets a uniform radius in [rmin, rmax] and a uniform random solid phase. Coordinates are interior
    cell indices (the reference adds the 3-cell halo offset itself, ``functions/fill_domain.h:75-76``).
    """
    pts: List[Tuple[int, int]] = []
    cell = min_spacing / math.sqrt(2.0)
    gx, gy = int(math.ceil(mesh_x / cell)), int(math.ceil(mesh_y / cell))
    grid = -np.ones((gx, gy), dtype=np.int64)
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place %d seeds with spacing %g in %dx%d" % (n, min_spacing, mesh_x, mesh_y))
        x = int(rng.integers(margin, mesh_x - margin))
        y = int(rng.integers(margin, mesh_y - margin))
        ix, iy = int(x / cell), int(y / cell)
        ok = True
        for i in range(max(0, ix - 2), min(gx, ix + 3)):
            for j in range(max(0, iy - 2), min(gy, iy + 3)):
                k = grid[i, j]
                if k >= 0:
                    px, py = pts[k]
                    if (px - x) ** 2 + (py - y) ** 2 < min_spacing ** 2:
                        ok = False
                        break
            if not ok:
                break
        if ok:
            grid[ix, iy] = len(pts)
            pts.append((x, y))
    fills = []
    for (x, y) in pts:
        r = float(rng.uniform(rmin, rmax))
        ph = int(rng.integers(0, solid_phases))
        fills.append(Fill("cylinder", ph, (x, y, round(r, 3))))
    return fills


def random_spd_A(rng: np.random.Generator, nc: int, lo: float = 0.5, hi: float = 2.0) -> List[float]:
    """Random symmetric positive-definite Hessian with eigenvalues in [lo, hi], returned as the ``A`` key tuple.

    The parabolic free energy is f = sum_{i<=j} A_ij c_i c_j + ..., so its Hessian is H_ii = 2 A_ii and
    H_ij = A_ij (``functions/functionF_01.h:5-18``); the eigenvalue bounds apply to H.
    """
    ev = rng.uniform(lo, hi, size=nc)
    q, _ = np.linalg.qr(rng.standard_normal((nc, nc)))
    H = (q * ev) @ q.T
    H = 0.5 * (H + H.T)
    diag = [round(float(H[i, i]) / 2.0, 6) for i in range(nc)]
    off = [round(float(H[i, j]), 6) for i in range(nc) for j in range(i + 1, nc)]
    return diag + off


def upper(n: int, v: float) -> List[float]:
    return [v] * (n * (n - 1) // 2)


# ----------------------------------------------------------------------------------------------------------------
# Legacy VTK (binary big-endian / ASCII), interior cells, x outer, y inner
# ----------------------------------------------------------------------------------------------------------------


def write_vtk(path: str, phi: np.ndarray, mu: np.ndarray, phase_names: Sequence[str], comp_names: Sequence[str],
              dx: float = 1.0, dy: float = 1.0, dz: float = 1.0, binary: bool = True) -> None:
    """Write interior fields ``phi[N, MX, MY]`` and ``mu[NC, MX, MY]`` in the serial driver's VTK layout."""
    N, MX, MY = phi.shape
    with open(path, "wb") as f:
        hdr = ("# vtk DataFile Version 3.0\nMicrosim_fields\n%s\nDATASET STRUCTURED_POINTS\n"
               "DIMENSIONS %d %d 1\nORIGIN 0 0 0\nSPACING %s %s %s\nPOINT_DATA %d\n"
               % ("BINARY" if binary else "ASCII", MY, MX, _le(dx), _le(dy), _le(dz), MX * MY))
        f.write(hdr.encode())
        blocks = [(n, phi[i]) for i, n in enumerate(phase_names[:N])]
        blocks += [("Mu_%s" % comp_names[k], mu[k]) for k in range(mu.shape[0])]
        for name, arr in blocks:
            f.write(("SCALARS %s double 1\nLOOKUP_TABLE default\n" % name).encode())
            if binary:
                f.write(np.ascontiguousarray(arr, dtype=">f8").tobytes())
            else:
                f.write("".join("%s\n" % _le(v) for v in arr.ravel()).encode())
                f.write(b"\n")


def _le(v: float) -> str:
    """C's ``%le`` formatting."""
    return "%e" % v


def read_vtk_fields(path: str, mesh_x: int, mesh_y: int) -> Dict[str, np.ndarray]:
    """Every SCALARS block of a binary or ASCII VTK written by the serial driver, by name, as [MX, MY] arrays."""
    with open(path, "rb") as f:
        data = f.read()
    binary = data.split(b"\n")[2].strip() == b"BINARY"
    n = mesh_x * mesh_y
    out: Dict[str, np.ndarray] = {}
    pos = data.index(b"SCALARS")
    while True:
        e = data.index(b"\n", pos)
        name = data[pos:e].split()[1].decode()
        pos = data.index(b"\n", e + 1) + 1   # past LOOKUP_TABLE
        if binary:
            out[name] = np.frombuffer(data, dtype=">f8", count=n, offset=pos).astype(np.float64).reshape(mesh_x, mesh_y)
            pos += 8 * n
        else:
            vals = []
            while len(vals) < n:
                e = data.index(b"\n", pos)
                t = data[pos:e].strip()
                pos = e + 1
                if t:
                    vals.append(float(t))
            out[name] = np.array(vals).reshape(mesh_x, mesh_y)
        nxt = data.find(b"SCALARS", pos)
        if nxt < 0:
            return out
        pos = nxt


def read_vtk(path: str, nphases: int, ncomp: int, mesh_x: int, mesh_y: int, extra_fields: int = 0
             ) -> Tuple[np.ndarray, np.ndarray]:
    """Read a binary or ASCII VTK written by the serial driver; returns (phi[N,MX,MY], mu[NC,MX,MY])."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0

    def line():
        nonlocal pos
        e = data.index(b"\n", pos)
        s = data[pos:e]
        pos = e + 1
        return s

    hdr = [line() for _ in range(8)]
    binary = hdr[2].strip() == b"BINARY"
    n = mesh_x * mesh_y
    out = []
    for _ in range(nphases + ncomp - 1):
        s = line()
        while not binary and not s.strip():
            s = line()
        line()
        if binary:
            arr = np.frombuffer(data, dtype=">f8", count=n, offset=pos).astype(np.float64)
            pos += 8 * n
        else:
            vals = []
            while len(vals) < n:
                s = line().strip()
                if s:
                    vals.append(float(s))
            arr = np.array(vals)
            # blank separator line
        out.append(arr.reshape(mesh_x, mesh_y))
    phi = np.stack(out[:nphases])
    mu = np.stack(out[nphases:]) if ncomp > 1 else np.zeros((0, mesh_x, mesh_y))
    return phi, mu


# ----------------------------------------------------------------------------------------------------------------
# Workloads
# ----------------------------------------------------------------------------------------------------------------


def binary_material(inp: GPInput, c_s: float = 0.2, c_l: float = 0.4, c_inf: float = 0.34, A: float = 1.0,
                    D_s: float = 0.0, D_l: float = 1.0, slope: float = 0.45) -> GPInput:
    """Binary solid/liquid parabolic material; the liquid is filled at ``c_inf`` (supersaturation
    (c_l - c_inf)/(c_l - c_s))."""
    inp.A = {0: [A], 1: [A]}
    inp.ceq = {(0, 0): [c_s], (0, 1): [c_l], (1, 1): [c_l], (1, 0): [c_l]}
    inp.cfill = {(0, 0): [c_s], (0, 1): [c_l], (1, 1): [c_inf], (1, 0): [c_l]}
    inp.slopes = {(0, 0): [slope], (0, 1): [slope], (1, 0): [slope], (1, 1): [slope]}
    inp.diffusivity = {0: [D_s], 1: [D_l]}
    return inp


def hessian(a: Sequence[float], nc: int) -> np.ndarray:
    """Hessian of f = sum_{i<=j} A_ij c_i c_j from the ``A`` key tuple (diagonal entries, then i<j pairs)."""
    H = np.diag([2.0 * a[i] for i in range(nc)])
    k = nc
    for i in range(nc):
        for j in range(i + 1, nc):
            H[i, j] = H[j, i] = a[k]
            k += 1
    return H


def fill_for_driving_force(inp: GPInput, driving: float) -> np.ndarray:
    """Liquid fill composition c that gives every solid a the same linearised grand-potential driving force

        psi_L(mu) - psi_a(mu) ~= (c_a - c_aL) . (mu - mu_a,eq) = (c_a - c_aL)^T H_L (c - c_aL) = driving,

    with c_a = ceq[a][a], c_aL = ceq[a][L] the two-phase equilibrium pair (the reference builds B_a, C_a so
    that they share mu and psi at Teq, functions/functionF_01.h:62-104) and mu_L(c) = H_L c. With as many
    solids as components this is a square solve; otherwise the minimum-norm least-squares solution."""
    N, NC = inp.nphases, inp.ncomp - 1
    L = N - 1
    HL = hessian(inp.A[L], NC)
    R = np.array([HL @ (np.asarray(inp.ceq[(a, a)]) - np.asarray(inp.ceq[(a, L)])) for a in range(N - 1)])
    rhs = np.array([R[a] @ np.asarray(inp.ceq[(a, L)]) + driving for a in range(N - 1)])
    return np.linalg.lstsq(R, rhs, rcond=None)[0]


def ternary_material(inp: GPInput, rng: np.random.Generator, D_l: float = 1.0, D_s: float = 0.0,
                     supersat: float = 0.3, driving: Optional[float] = None) -> GPInput:
    """N-phase ternary material: random SPD Hessians (eigenvalues in [0.5, 2]) and equilibrium compositions
    uniform in [0.05, 0.45]. The liquid is filled between its equilibrium compositions and the solids'
    (``supersat``), or, when ``driving`` is given, at the composition that gives every solid that linearised
    driving force (fill_for_driving_force)."""
    N, NC = inp.nphases, inp.ncomp - 1
    L = N - 1
    inp.A = {p: random_spd_A(rng, NC) for p in range(N)}
    ceq, cfill, slopes = {}, {}, {}
    cl_mean = np.zeros(NC)
    for a in range(N - 1):
        cs = np.round(rng.uniform(0.05, 0.45, NC), 4)
        cl = np.round(rng.uniform(0.05, 0.45, NC), 4)
        ceq[(a, a)] = list(cs)
        ceq[(a, L)] = list(cl)
        ceq[(L, a)] = list(cl)
        cfill[(a, a)] = list(cs)
        cfill[(a, L)] = list(cl)
        cfill[(L, a)] = list(cl)
        slopes[(a, a)] = [0.45] * NC
        slopes[(a, L)] = [0.45] * NC
        slopes[(L, a)] = [0.45] * NC
        cl_mean += cl
    cl_mean /= (N - 1)
    cs_mean = np.mean([ceq[(a, a)] for a in range(N - 1)], axis=0)
    ceq[(L, L)] = list(np.round(cl_mean, 4))
    cfill[(L, L)] = list(np.round(cl_mean - supersat * (cl_mean - cs_mean), 4))
    slopes[(L, L)] = [0.45] * NC
    inp.ceq, inp.cfill, inp.slopes = ceq, cfill, slopes
    inp.diffusivity = {p: [D_s] * NC for p in range(N - 1)}
    inp.diffusivity[L] = [D_l] * NC
    if driving is not None:
        cf = fill_for_driving_force(inp, driving)
        if not np.all((cf > 0.0) & (cf < 1.0)) or cf.sum() >= 1.0:
            raise ValueError("driving force %g needs an unphysical liquid fill %s" % (driving, cf))
        cfill[(L, L)] = list(np.round(cf, 4))
    return inp


def tanh_seed_state(mesh_x: int, mesh_y: int, radius: float, width: float, mu_liquid: Sequence[float],
                    nphases: int = 2, centre: Optional[Tuple[float, float]] = None) -> Tuple[np.ndarray, np.ndarray]:
    """Circular solid seed (phase 0) with a tanh profile, liquid (last phase) elsewhere, uniform mu."""
    cx, cy = centre if centre is not None else ((mesh_x - 1) / 2.0, (mesh_y - 1) / 2.0)
    x = np.arange(mesh_x, dtype=np.float64)[:, None]
    y = np.arange(mesh_y, dtype=np.float64)[None, :]
    r = np.sqrt((x - cx) ** 2 + (y - cy) ** 2)
    ps = 0.5 * (1.0 - np.tanh((r - radius) / width))
    phi = np.zeros((nphases, mesh_x, mesh_y))
    phi[0] = ps
    phi[nphases - 1] = 1.0 - ps
    mu = np.stack([np.full((mesh_x, mesh_y), float(m)) for m in mu_liquid])
    return phi, mu


def config_c0(mesh: int = 200, steps: int = 20000, savet: int = 2000) -> Tuple[GPInput, List[Fill]]:
    """Binary two-phase anisotropic seed: 200x200, dx=1, dt=0.05, eps=4, gamma=1, tau=0.3, D_l=1, D_s=0,
    dab=0.02, c_s=0.2, c_l=0.4, liquid at 0.34 (supersaturation 0.3), Neumann, tanh seed r=10 (restart file)."""
    inp = GPInput(mesh_x=mesh, mesh_y=mesh, nphases=2, ncomp=2, dx=1.0, dy=1.0, dt=0.05, ntimesteps=steps,
                  savet=savet, nsmooth=1, epsilon=4.0, tau=0.3, gamma=[1.0], tau_ab=[0.3], dab=[0.02],
                  components=["Al", "Cu"], phases=["alpha", "liquid"], bc_phi=(1, 1, 1, 1), restart=1,
                  track_progress=max(1, savet))
    binary_material(inp)
    return inp, []


def config_ternary(mesh: int, nseeds: int, seed: int = 1234, steps: int = 1000, savet: int = 1000,
                   dt_frac: float = 0.8, min_spacing: float = 60.0, gamma: float = 0.2,
                   driving: Optional[float] = 0.04) -> Tuple[GPInput, List[Fill], float]:
    """Ternary three-phase random-seed workload (two solids + liquid), periodic boundaries, dt at ``dt_frac``
    of the explicit stability limit. Returns (input, fills, dt_limit).

    Interface energy and supersaturation are chosen so that the seeds (radius 6-12) are supercritical and keep
    growing: the liquid is filled at a driving force of 0.04 for both solids (critical radius gamma/driving =
    5 cells at gamma = 0.2). With gamma = 1 and compositions in [0.05, 0.45], every seed dissolves within about
    a thousand steps and the run degenerates into pure diffusion. ``driving=None`` gives the older
    supersaturation-0.3 fill."""
    rng = np.random.default_rng(seed)
    inp = GPInput(mesh_x=mesh, mesh_y=mesh, nphases=3, ncomp=3, dx=1.0, dy=1.0, ntimesteps=steps, savet=savet,
                  nsmooth=1, epsilon=4.0, tau=0.3, gamma=upper(3, gamma), tau_ab=upper(3, 0.3),
                  dab=upper(3, 0.02), gamma_abc=[0.0], components=["Al", "Cu", "Ni"],
                  phases=["alpha", "beta", "liquid"], bc_phi=(3, 3, 3, 3), track_progress=max(1, savet))
    ternary_material(inp, rng, driving=driving)
    dtl = explicit_dt_limit(inp)
    inp.dt = float("%.6g" % (dt_frac * dtl))
    fills = poisson_disc_seeds(rng, mesh, mesh, nseeds, 6.0, 12.0, min_spacing, 2, margin=13)
    return inp, fills, dtl


def explicit_dt_limit(inp: GPInput) -> float:
    """Explicit-Euler stability bound of the two coupled equations on a uniform grid.

    mu-equation: FTCS diffusion with the largest phase diffusivity, dt <= 1 / (2 D (1/dx^2 + 1/dy^2)).
    phi-equation: for a two-phase interface q_ab = -grad(phi_a), the Lagrange-mean form gives
    dphi/dt = (Gamma/tau) lap(phi) (isotropic), so dt <= tau / (2 Gamma_eff (1/dx^2 + 1/dy^2)) with
    Gamma_eff = Gamma_max (1 + 15 dab_max) (4-fold stiffness bound) times 2(N-1)/N for the N-phase mean.
    """
    D = max(max(v) for v in inp.diffusivity.values())
    lap = 1.0 / inp.dx ** 2 + 1.0 / inp.dy ** 2
    dt_mu = 1.0 / (2.0 * D * lap) if D > 0 else float("inf")
    g = max(inp.gamma) if inp.gamma else 1.0
    d = max(inp.dab) if (inp.dab and inp.fn_anisotropy) else 0.0
    tau_min = min(list(inp.tau_ab) + [inp.tau])
    geff = g * (1.0 + 15.0 * d) * 2.0 * (inp.nphases - 1) / inp.nphases
    dt_phi = tau_min / (2.0 * geff * lap)
    return min(dt_mu, dt_phi)


def write_case(dirname: str, inp: GPInput, fills: Sequence[Fill], restart_state=None) -> None:
    """Write Input.in / Filling.in (and DATA/<out>_0.vtk for a restart state) into ``dirname``."""
    os.makedirs(os.path.join(dirname, "DATA"), exist_ok=True)
    with open(os.path.join(dirname, "Input.in"), "w") as f:
        f.write(inp.text())
    with open(os.path.join(dirname, "Filling.in"), "w") as f:
        f.write(filling_text(fills) if fills else "# no fills\n")
    if restart_state is not None:
        phi, mu, out = restart_state
        comps = list(inp.components) or ["C%d" % i for i in range(inp.ncomp)]
        phases = list(inp.phases) or (["S%d" % i for i in range(inp.nphases - 1)] + ["L"])
        write_vtk(os.path.join(dirname, "DATA", "%s_%d.vtk" % (out, inp.starttime)), phi, mu, phases, comps,
                  inp.dx, inp.dy, inp.dx, binary=(inp.writeformat == "BINARY"))
