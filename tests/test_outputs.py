"""Outputs and restart: FARGO3D's merged raw files written from device state,
and a run restarted from them.

The layout (WriteMerging, src/output.c:283-315: active cells, x fastest,
then y, then z, native doubles, no header) is pinned on host arrays whose
every cell holds a distinct value, so a transposed or shifted loop fails.
The GPU tests check the files against the device fields, the planet file
against the planet state to the last bit (`%#.18g` round-trips a double),
and that a run restarted from output n (src/main.c:249-260) continues bit
for bit like the run that wrote it.
"""
import ctypes
import os

import numpy as np
import pytest

C1 = "fargo3d_amd/setups/c1_fargo2d.par"
C2 = "fargo3d_amd/setups/c2_p3d.par"
C3 = "fargo3d_amd/setups/c3_adiabatic2d.par"
C4 = "fargo3d_amd/setups/c4_multifluid2d.par"


@pytest.fixture(scope="module")
def fa():
    import fargo3d_amd
    return fargo3d_amd


def _grid(nx, ny, nz, ngh, dimz):
    sy = ny + 2 * ngh
    sz = nz + 2 * ngh if dimz else 1
    k, j, i = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(nx), indexing="ij")
    return (i + 1000.0 * j + 1.0e6 * k).astype(np.float64)


@pytest.mark.parametrize("dimz", [0, 1])
def test_active_file_layout(fa, tmp_path, dimz):
    L = fa.lib()
    nx, ny, nz, ngh = 7, 5, 4, 3
    h = _grid(nx, ny, nz, ngh, dimz)
    sz, sy, _ = h.shape
    nzf, nghz = (nz, ngh) if dimz else (1, 0)
    path = str(tmp_path / "f.dat").encode()
    assert L.f3d_write_active(h.ctypes.data, nx, ny, nzf, ngh, nghz, nx, nx * sy, path) == 0
    raw = np.fromfile(tmp_path / "f.dat", dtype=np.float64)
    act = h[ngh:ngh + nz, ngh:ngh + ny, :] if dimz else h[:, ngh:ngh + ny, :]
    assert raw.size == nx * ny * nzf
    # x fastest, then y, then z: value = i + 1000 j + 1e6 k
    kk, jj, ii = np.unravel_index(np.arange(raw.size), (nzf, ny, nx))
    assert np.array_equal(raw, ii + 1000.0 * (jj + ngh) + 1.0e6 * (kk + nghz))
    assert np.array_equal(raw, act.ravel())
    # read-back fills the active cells and leaves the ghosts
    back = np.full_like(h, -1.0)
    assert L.f3d_read_active(back.ctypes.data, nx, ny, nzf, ngh, nghz, nx, nx * sy, path) == 0
    mask = np.zeros(h.shape, bool)
    if dimz:
        mask[ngh:ngh + nz, ngh:ngh + ny, :] = True
    else:
        mask[:, ngh:ngh + ny, :] = True
    assert np.array_equal(back[mask], h[mask])
    assert np.all(back[~mask] == -1.0)
    # a file of another size is refused (RestartDat would read a stale output)
    raw[:-1].tofile(tmp_path / "short.dat")
    assert L.f3d_read_active(back.ctypes.data, nx, ny, nzf, ngh, nghz, nx, nx * sy,
                             str(tmp_path / "short.dat").encode()) == -1
    assert L.f3d_read_active(back.ctypes.data, nx, ny, nzf, ngh, nghz, nx, nx * sy,
                             str(tmp_path / "none.dat").encode()) == -1


def _active(s, x):
    return x[3:-3, 3:-3, :] if s.dimz else x[:, 3:-3, :]


def _fluid_files(s, k):
    name = "gas" if k == 0 else "dust%d" % k
    q = [("dens", "Density")]
    if k == 0:
        q.append(("energy", "Energy"))
    q += [("vx", "Vx"), ("vy", "Vy")] + ([("vz", "Vz")] if s.dimz else [])
    return name, q


def _state(s):
    out = {}
    for k in range(s.nfluids):
        s.select_fluid(k)
        for _, f in _fluid_files(s, k)[1]:
            out[(k, f)] = _active(s, s.field(f))
    s.select_fluid(0)
    return out


CASES = {
    "c1_2d_iso": (C1, "Nx=64 Ny=32 Ntot=2 Ninterm=1"),
    "c2_3d_iso": (C2, "Nx=64 Ny=24 Nz=12 Ntot=2 Ninterm=1"),
    "c3_2d_adiabatic": (C3, "Nx=128 Ny=48 Ntot=2 Ninterm=1"),
    "c4_multifluid": (C4, "Nx=128 Ny=64 Ntot=2 Ninterm=1"),
    "c2_3d_precision_form": (C2, "Nx=64 Ny=24 Nz=12 Ntot=2 Ninterm=1 Precision=32"),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_outputs_and_restart(fa, tmp_path, case):
    par, ov = CASES[case]
    print("outputs/restart", case, flush=True)
    d = str(tmp_path)
    # the run that writes outputs 0, 1, 2 (NTOT = 2 passes of DT)
    m = fa.Simulation(par, ov)
    m.write_init(d)
    m.main_loop(0, d)
    final = _state(m)
    planets = m.planets()
    sc = m.scalars()
    nfl = m.nfluids
    # files of output 2 hold the device state
    for k in range(nfl):
        name, q = _fluid_files(m, k)
        for suf, f in q:
            raw = np.fromfile(os.path.join(d, "%s%s2.dat" % (name, suf)), dtype=np.float64)
            assert np.array_equal(raw, final[(k, f)].ravel()), (name, suf)
        if k > 0:
            assert not os.path.exists(os.path.join(d, "%senergy2.dat" % name))
    if not m.dimz:
        assert not os.path.exists(os.path.join(d, "gasvz2.dat"))
    # planet<k>.dat: lines 0, 1, 2; the last one is the state to the bit
    for p in range(planets.shape[0]):
        lines = open(os.path.join(d, "planet%d.dat" % p)).read().splitlines()
        assert [int(x.split("\t")[0]) for x in lines] == [0, 1, 2]
        v = [float(x) for x in lines[-1].split("\t")[1:]]
        assert v[:7] == list(planets[p]), p
        assert v[7] == sc["PhysicalTime"] and v[8] == sc["OMEGAFRAME"]
    m.close()

    # a fresh run restarted from output 1 continues like the run that wrote it
    r = fa.Simulation(par, ov)
    begin = r.restart(1, d)
    assert begin == 1
    nsteps = r.advance()
    assert nsteps > 0
    got = _state(r)
    worst = 0
    for key, ref in final.items():
        nb = int(np.count_nonzero(got[key].view(np.int64) != ref.view(np.int64)))
        worst = max(worst, nb)
        print("  ", key, "bits differ:", nb, flush=True)
    assert worst == 0
    rp = r.planets()
    assert np.array_equal(rp, planets)
    assert r.scalars()["PhysicalTime"] == sc["PhysicalTime"]
    r.close()


@pytest.mark.gpu
def test_restart_refuses_missing_output(fa, tmp_path):
    s = fa.Simulation(C1, "Nx=64 Ny=32")
    with pytest.raises(fa.HipError, match="cannot restart"):
        s.restart(3, str(tmp_path))
    s.close()
