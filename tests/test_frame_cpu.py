"""Host side of the non-isothermal runs and the moving frame, against the reference's own runs of the frame decks
(tests/golden: bin_tempgrad_shift, a planar front under a moving temperature gradient whose frame shifts by two
rows at step 100; tern_tempgrad, three-phase seeds under a gradient).

- gp_temperature_rows restates apply_temperature_gradientY (Grand_potential_Finite_difference_2D_serial/
  functions/Temperature_gradient.h:4-29) with the BASE_POS / GRADIENT / temp_bottom of microsim_gp.c:77-80 and
  108-111; the reference's Temperature blocks pin it at every stored step, including the rows a shift moved.
- The input stage derives the frame parameters the loop reads.
"""
import os

import numpy as np
import pytest

from gp_testlib import GOLDEN, frame_cases, load_golden


def _shift_positions(name):
    out = {}
    for line in open(os.path.join(GOLDEN, name, "shift.dat")):
        t, pos = line.split()
        out[int(t)] = int(pos)
    return out


def test_frame_decks_present():
    assert set(frame_cases()) == {"bin_tempgrad_shift", "tern_tempgrad"}


@pytest.mark.parametrize("name", frame_cases())
def test_temperature_rows_match_reference_blocks(name):
    from microsim_amd import host
    case, ref = load_golden(name)
    assert "T" in ref
    pos = _shift_positions(name)
    p = case.params
    my = int(p.mesh_y)
    for i, s in enumerate(ref["steps"]):
        want = ref["T"][i]
        assert np.all(want == want[:1]), "T depends on y only"
        off_before = pos.get(s - 1, 0) if s > 0 else 0
        off = pos.get(s, 0)
        if s > 0 and s % 100 == 0 and off != off_before:
            # the frame moved after this step's gradient: rows y <= rows_y - 1 - k take row y + k, the top k keep
            # their own (apply_shiftY, functions/shift.h:4-49)
            k = off - off_before
            T = host.temperature_rows(p, s, off_before)
            T[:len(T) - k] = T[k:].copy()
        else:
            T = host.temperature_rows(p, s, off)
        assert np.array_equal(T[3:my + 3], want[0]), "%s step %d: max diff %g" % (
            name, s, np.abs(T[3:my + 3] - want[0]).max())


def test_shifted_deck_moves_once():
    pos = _shift_positions("bin_tempgrad_shift")
    assert all(v == 0 for t, v in pos.items() if t < 100) and all(v == 2 for t, v in pos.items() if t >= 100)


def test_frame_parameters():
    case, _ = load_golden("bin_tempgrad_shift")
    p = case.params
    assert p.tempgrady == 1 and p.shift == 1 and p.shiftj == 16
    assert (p.tg_base_temp, p.tg_deltaT, p.tg_distance, p.tg_offset, p.tg_velocity) == (0.97, 0.02, 40.0, 15.0, 0.5)
    # the liquid rows a shift adds hold Mu(cfill[L][L], Tfill, L) (functions/shift.h:24-38): 2 A c + B(Tfill)
    L = int(p.nphases) - 1
    c = case.inp.cfill[L][L][0]
    assert c == 0.34
    assert p.mu_fill[0] == 2.0 * p.A[(L * 4 + 0) * 4 + 0] * c + (p.Beq[L * 4] + p.dBbdT[L * 4] * (1.0 - p.Teq))
