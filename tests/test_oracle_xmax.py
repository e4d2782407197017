"""The x-max Silver-Mueller boundary (radiating_boundary, boundary.cpp:92-96)
and the face form PIC_OPT_XMAX_BOUNDARY = 1 selects.

The reference writes B at column local_xcells+1 from E, Bx, By, Bz and J at
column local_xcells, while its x-min write (boundary.cpp:87-91) reads the
primal column next to the dual column it writes.  Read one column too far
in, the x-max condition is unstable once a wave reaches it: a vacuum pulse
that leaves through x-max drives the field energy up without bound, at
either order.  The face form mirrors the x-min write at the last dual
column and absorbs the pulse.  Both are checked here on the oracle (no
GPU), through a pulse that crosses a 64-cell box and leaves it.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from pic3d_amd import workloads as W


def vacuum_run(mem, face, blocks=13, nthreads=4):
    wl = W.laser_slab(nx=64, ny=8, nz=8, ppc=1, slab=(10.0, 11.0), mem=mem)
    o = Oracle(wl.params, nthreads)
    o.set_xmax_face(face)
    o.species_set(0, -1.0, 1.0)
    o.species_set(1, 1.0, 1836.0, True)
    energy = []
    for blk in range(blocks):
        o.step(1 + 50 * blk, 50)
        energy.append(o.diag(1.0)["field_energy"])
    return o, np.array(energy)


@pytest.mark.parametrize("mem", [2, 4])
def test_face_form_absorbs_reference_form_grows(mem):
    _, face = vacuum_run(mem, True)
    _, ref = vacuum_run(mem, False)
    peak = face.max()
    assert face[-1] < 1e-5 * peak, face        # the 10-cycle pulse has left the box
    assert ref[-1] > 10 * ref[:8].max(), ref   # the reference's condition grows
    # the same until the pulse reaches x-max (0.5 cells per step: about step 128)
    assert np.array_equal(face[:2], ref[:2])
    assert not np.allclose(face[3:], ref[3:], rtol=1e-6)


def test_forms_identical_away_from_xmax():
    fields = []
    for face in (False, True):
        o, _ = vacuum_run(4, face, blocks=2, nthreads=1)
        fields.append([o.field_get(c) for c in range(6)])
    m, n = 4, 64
    for a, b in zip(*fields):
        # only the x-max write column and what it reached in 100 steps differ
        assert np.array_equal(a[..., :n - 20], b[..., :n - 20])
