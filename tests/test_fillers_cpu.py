"""The MPI driver's random and pattern fillers in the host's fill_domain (gp_fill_domain, microsim_amd/host/gp_host.c),
against restatements of Grand_potential_Finite_difference_2D_MPI/functions/filling.h on the same random streams.

- FILLCYLINDERRANDOM, FILLSPHERERANDOM and FILLVORONOI2D draw from drand48/lrand48, which the reference never seeds:
  every run uses the POSIX initial state. The restatements here draw from the C library's own erand48/nrand48 from
  that state, so the host's generator and its use are pinned exactly (phi equal cell for cell).
- FILLCUBEPATTERN and FILLCUBERANDOM draw from GSL's default MT19937 seeded with time(0) in the reference; the host
  takes GP_FILL_SEED when set. The restatements draw the raw 32-bit outputs of numpy's legacy MT19937 with the same
  seed (its integer seeding is the same Knuth recurrence as GSL's), so these are pinned exactly as well, and by
  structure: particle counts, sizes and phases.
"""
import ctypes as C
import math
import os
import shutil
import tempfile

import numpy as np
import pytest

from gp_testlib import ensure_host_built
from microsim_amd import _abi, configs, host

_libc = C.CDLL("libc.so.6")
_libc.erand48.argtypes = [C.POINTER(C.c_ushort)]
_libc.erand48.restype = C.c_double
_libc.nrand48.argtypes = [C.POINTER(C.c_ushort)]
_libc.nrand48.restype = C.c_long

MESH = 64
_DIRS = []


@pytest.fixture(autouse=True, scope="module")
def _cleanup():
    yield
    for d in _DIRS:
        shutil.rmtree(d, ignore_errors=True)


class Drand48:
    """erand48 / nrand48 of the C library from the POSIX initial state 0x1234ABCD330E."""

    def __init__(self):
        self.x = (C.c_ushort * 3)(0x330E, 0xABCD, 0x1234)

    def drand(self):
        return _libc.erand48(self.x)

    def lrand(self):
        return _libc.nrand48(self.x)


class MT:
    """GSL-style uniform draws from numpy's legacy MT19937 with an integer seed."""

    def __init__(self, seed):
        self.rs = np.random.RandomState(seed)

    def raw(self):
        return int(self.rs.randint(0, 2 ** 32, dtype=np.uint32))

    def uniform(self):
        return self.raw() / 4294967296.0

    def uniform_int(self, n):
        scale = 0xffffffff // n
        while True:
            k = self.raw() // scale
            if k < n:
                return k


def _case(nphases=3):
    if nphases == 3:
        inp, fills, _ = configs.config_ternary(MESH, 1, seed=7, min_spacing=20.0)
    else:
        inp, fills = configs.config_c0(MESH, steps=10, savet=10)
    d = tempfile.mkdtemp(prefix="gp_fill_")
    _DIRS.append(d)
    configs.write_case(d, inp, fills)
    return d, host.read_input(os.path.join(d, "Input.in"))


def _fill(d, inp, lines, seed=None):
    ensure_host_built()
    path = os.path.join(d, "Filling.in")
    open(path, "w").write("".join(l + "\n" for l in lines))
    rx, ry = MESH + 6, MESH + 6
    N, NC = int(inp.p.nphases), int(inp.p.ncomp) - 1
    phi, mu = np.zeros((N, rx, ry)), np.zeros((NC, rx, ry))
    err = C.create_string_buffer(512)
    old = os.environ.get("GP_FILL_SEED")
    if seed is not None:
        os.environ["GP_FILL_SEED"] = str(seed)
    try:
        rc = _abi.host_lib().gp_fill_domain(C.byref(inp), path.encode(), _abi.dptr(phi), _abi.dptr(mu), None, 1, err,
                                            512)
    finally:
        if seed is not None:
            if old is None:
                del os.environ["GP_FILL_SEED"]
            else:
                os.environ["GP_FILL_SEED"] = old
    assert rc == 0, err.value
    return phi, mu


# ------------------------------------------------------------------ restatements (filling.h of the MPI driver)
def _solid(phi, b, inside):
    """fill_phase_cube / *_occupancy for a solid b over the whole grid."""
    N = phi.shape[0]
    for a in range(N):
        if a != b:
            phi[a][inside] = 0.0
    outside = ~inside & (phi[b] != 1.0)
    phi[b][inside] = 1.0
    phi[b][outside] = 0.0


def _liquid(phi):
    N = phi.shape[0]
    s = np.zeros(phi.shape[1:])
    for a in range(N - 1):
        s += phi[a]
    wrong = s > 1.0
    phi[N - 1] = np.where(wrong, 1.0, 1.0 - s)
    for a in range(N - 1):
        phi[a][wrong] = 0.0


def _grid():
    x, y = np.meshgrid(np.arange(MESH + 6), np.arange(MESH + 6), indexing="ij")
    return x, y


def _cylinders(phi, r, phase, rad0, vf, shield, spread, sphere=False):
    shield = 8 if shield > 8 else (2 if shield == 1 else shield)
    x, y = _grid()
    occ = np.zeros(x.shape, dtype=bool)
    vol = float(MESH) * MESH * 1
    vpp = float(rad0 * rad0 * rad0) * (4.0 / 3.0) * math.pi if sphere else float(rad0 * rad0) * math.pi
    n = int(math.ceil(vol * vf / vpp))
    idx, count, limit = 1, 0, (100000 if sphere else 10000)
    while idx <= n:
        cx, cy = int(MESH * r.drand()), int(MESH * r.drand())
        cz = int(1 * r.drand()) if sphere else 0
        rad = float(rad0) + (1.0 * r.drand() - 0.5) * (spread * rad0)
        count += 1
        if count > limit:
            break
        if cx - rad < 0.0 or cx + rad > MESH - 1 or cy - rad < 0.0 or cy + rad > MESH - 1:
            continue
        if sphere and (cz - rad < 0.0 or cz + rad > 0):
            continue
        d2 = (x - cx) ** 2 + (y - cy) ** 2 + (0 - cz) ** 2
        if shield and np.any(occ & (d2 <= (shield * rad) * (shield * rad))):
            continue
        inside = d2 <= rad * rad
        _solid(phi, phase, inside)
        occ |= inside
        idx += 1
    _liquid(phi)
    return idx - 1


def _voronoi(phi, r, x0, x1, y0, y1, npts, size_min):
    N = phi.shape[0]
    flag = np.zeros(phi.shape[1:], dtype=bool)
    pts = []
    for _ in range(npts):
        rx = int(x0 + r.drand() * (x1 - x0))
        ry = int(y0 + r.drand() * (y1 - y0))
        if not flag[rx, ry]:
            pts.append((float(rx), float(ry)))
            for x in range(x0, x1):
                for y in range(y0, y1):
                    if (ry - y) ** 2 + (rx - x) ** 2 <= size_min * size_min:
                        flag[x, y] = True
        else:
            pts.append(None)
    ph = [r.lrand() % (N - 1) for _ in range(npts)]
    for x in range(x0, x1):
        for y in range(y0, y1):
            minimum, location = (pts[0][0] - x) ** 2 + (pts[0][1] - y) ** 2, 0
            for k in range(1, npts):
                if pts[k] is None:
                    continue
                l = (pts[k][0] - x) ** 2 + (pts[k][1] - y) ** 2
                if l < minimum:
                    minimum, location = l, ph[k]
            for s in range(N - 1):
                phi[s, x, y] = 1.0 if location == s else 0.0
    _liquid(phi)


def _cube(phi, b, x0, x1, y0, y1, z0, z1):
    x, y = _grid()
    _solid(phi, b, (x >= x0) & (x <= x1) & (y >= y0) & (y <= y1) & (0 >= z0) & (0 <= z1))


def _pattern(phi, mt, variants, sx, sy, sfrac, gap, gfrac, outof):
    N = phi.shape[0]
    phi[N - 1] = 1.0
    qx, qy = max(MESH // (sx + gap), 1), max(MESH // (sy + gap), 1)
    placed = []
    xlo = ylo = 0
    xhi = yhi = 1
    for i in range(qx):
        r = mt.uniform()
        sgn = 2 * math.floor(r + 0.5) - 1
        xlo = int(i * (gap + sx) + gap + sgn * gfrac * gap - (2 * r - 1) * sfrac * sx)
        xhi = int((i + 1) * (gap + sx) + sgn * gfrac * gap + (2 * r - 1) * sfrac * sx)
        for j in range(qy):
            r = mt.uniform()
            sgn = 2 * math.floor(r + 0.5) - 1
            ylo = int(j * (gap + sy) + gap + sgn * gfrac * gap - (2 * r - 1) * sfrac * sy)
            yhi = int((j + 1) * (gap + sy) + sgn * gfrac * gap + (2 * r - 1) * sfrac * sy)
            if outof > 0 and not mt.uniform_int(outof):
                continue
            variants = min(variants, N - 1)
            phase = int(mt.uniform() * variants)
            _cube(phi, phase, xlo, xhi, ylo, yhi, 0, 1)
            placed.append((phase, xlo, xhi, ylo, yhi))
    return placed


def _random_cubes(phi, mt, variants, sx, sy, sz, sfrac, vf, shield):
    N = phi.shape[0]
    ry = phi.shape[2]
    phi[N - 1] = 1.0
    vol = MESH * MESH * 1
    n = int(math.ceil(vol * vf / float(sx * sy * sz)))
    placed, count = [], 0
    flat = phi[N - 1].reshape(-1)
    while n:
        r = mt.uniform()
        sxn, syn, szn = (int(s + (2 * r - 1) * sfrac * s) for s in (sx, sy, sz))
        xn, yn, zn = int(MESH * mt.uniform()), int(MESH * mt.uniform()), int(1 * mt.uniform())
        c = lambda v, s: (v - int(s / 2), v + int(s / 2))
        (xlo, xhi), (ylo, yhi), (zlo, zhi) = c(xn, sxn), c(yn, syn), c(zn, szn)
        overlap = False
        for x in range(xlo - shield, xhi + shield):
            for y in range(ylo - shield, yhi + shield):
                for z in range(zlo - shield, zhi + shield):
                    index = x * ry + z * ry + y
                    if 0 <= index <= vol and not (flat[index] > 0):
                        overlap = True
                        break
                if overlap:
                    break
            if overlap:
                break
        if overlap:
            continue
        variants = min(variants, N - 1)
        phase = int(mt.uniform() * variants)
        _cube(phi, phase, xlo, xhi, ylo, yhi, zlo, zhi)
        placed.append((phase, xlo, xhi, ylo, yhi))
        n -= 1
        count += 1
        if count > 10000:
            break
    return placed


# ------------------------------------------------------------------ tests
def test_drand48_restatement_is_the_c_library():
    """The POSIX recurrence and initial state, stated independently, give the C library's stream (drand48 = X / 2^48,
    lrand48 = X >> 17)."""
    r, x = Drand48(), 0x1234ABCD330E
    for i in range(1000):
        x = (0x5DEECE66D * x + 0xB) % (1 << 48)
        if i % 2:
            assert r.lrand() == x >> 17
        else:
            assert r.drand() == x / 2.0 ** 48


def test_mt19937_restatement():
    """Known first outputs of MT19937 for seed 5489, and GSL's uniform scaling."""
    mt = MT(5489)
    assert [mt.raw() for _ in range(3)] == [3499211612, 581869302, 3890346734]


def test_cylinders_random_match_restatement():
    d, inp = _case()
    line = "FILLCYLINDERRANDOM = {1, 5, 0.15, 2, 0.4};"
    phi, mu = _fill(d, inp, [line])
    want = np.zeros_like(phi)
    n = _cylinders(want, Drand48(), 1, 5, 0.15, 2, 0.4)
    assert n >= 5
    assert np.array_equal(phi, want)
    # structure: only phase 1 and the liquid, each cell pure
    assert np.all(phi[0] == 0.0) and np.all((phi[1] == 0.0) | (phi[1] == 1.0))
    assert np.array_equal(phi[1] + phi[2], np.ones_like(phi[1]))


def test_spheres_random_place_nothing_in_2d_but_draw():
    """FILLSPHERERANDOM on a one-cell-deep grid fails its z-extent test on every attempt; it still draws four values
    per attempt up to its 1e5 limit, which a following random filler sees."""
    d, inp = _case()
    lines = ["FILLSPHERERANDOM = {0, 4, 0.1, 0, 0.2};", "FILLCYLINDERRANDOM = {0, 4, 0.1, 0, 0.2};"]
    phi, _ = _fill(d, inp, lines)
    want = np.zeros_like(phi)
    r = Drand48()
    assert _cylinders(want, r, 0, 4, 0.1, 0, 0.2, sphere=True) == 0
    assert _cylinders(want, r, 0, 4, 0.1, 0, 0.2) > 0
    assert np.array_equal(phi, want)


def test_voronoi_matches_restatement():
    d, inp = _case()
    phi, _ = _fill(d, inp, ["FILLVORONOI2D = {0, 64, 0, 64, 12, 6.0};"])
    want = np.zeros_like(phi)
    _voronoi(want, Drand48(), 3, 67, 3, 67, 12, 6.0)
    assert np.array_equal(phi, want)
    inner = phi[:, 3:67, 3:67]
    assert np.all(inner[2] == 0.0) and np.array_equal(inner[0] + inner[1], np.ones_like(inner[0]))
    assert inner[0].sum() > 0 and inner[1].sum() > 0


@pytest.mark.parametrize("seed", [11, 12345])
def test_cube_pattern_matches_restatement(seed):
    d, inp = _case()
    phi, _ = _fill(d, inp, ["FILLCUBEPATTERN = {2, 8, 8, 1, 0.2, 4, 0.3, 3};"], seed=seed)
    want = np.zeros_like(phi)
    placed = _pattern(want, MT(seed), 2, 8, 8, 0.2, 4, 0.3, 3)
    assert np.array_equal(phi, want)
    # structure: a 5x5 lattice (64 // 12), each box kept with probability 2/3, of phase 0 or 1
    assert 0 < len(placed) <= 25 and {p[0] for p in placed} <= {0, 1}
    for _, x0, x1, y0, y1 in placed:
        assert 8 * 0.8 - 3 <= x1 - x0 <= 8 * 1.2 + 5


@pytest.mark.parametrize("seed", [3, 999])
def test_random_cubes_match_restatement(seed):
    d, inp = _case()
    phi, _ = _fill(d, inp, ["FILLCUBERANDOM = {2, 6, 6, 1, 0.25, 0.12, 2};"], seed=seed)
    want = np.zeros_like(phi)
    placed = _random_cubes(want, MT(seed), 2, 6, 6, 1, 0.25, 0.12, 2)
    assert np.array_equal(phi, want)
    assert len(placed) == math.ceil(64 * 64 * 0.12 / 36.0)
    assert {p[0] for p in placed} <= {0, 1}


def test_fillers_reject_bad_tuples():
    d, inp = _case()
    path = os.path.join(d, "Filling.in")
    for line in ["FILLCYLINDERRANDOM = {2, 5, 0.1, 2, 0.4};", "FILLVORONOI2D = {0, 200, 0, 64, 4, 2.0};",
                 "FILLCUBERANDOM = {2, 0, 6, 1, 0.2, 0.1, 1};"]:
        open(path, "w").write(line + "\n")
        rx = ry = MESH + 6
        phi, mu = np.zeros((3, rx, ry)), np.zeros((2, rx, ry))
        err = C.create_string_buffer(512)
        rc = _abi.host_lib().gp_fill_domain(C.byref(inp), path.encode(), _abi.dptr(phi), _abi.dptr(mu), None, 1, err,
                                            512)
        assert rc != 0 and err.value, line
