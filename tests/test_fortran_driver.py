"""The Fortran ISO_C_BINDING driver (d3q19-single-phase_amd/fortran) end to end.

CPU: the checkpoint record format round-trips, and the driver binary exists
and links against liblbm.  GPU: the driver, run on the device through the
reference-named shim (collision_MRT, macrovar, ...), reproduces the oracle bit
for bit on an injected start (loadinitflow) and on a fresh start with the
pre-relaxation loop, writing its result in the reference's checkpoint format.
"""
import os
import subprocess

import numpy as np
import pytest

import d3q19
from d3q19 import fortio, workloads
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "d3q19-single-phase_amd", "fortran", "channel_main")


def test_record_format_roundtrip(tmp_path):
    f = workloads.uniform(4, 19 * 5 * 4 * 3).reshape(3, 4, 5, 19)
    p = tmp_path / "finit.000"
    fortio.write_finit(str(p), f, istat=7)
    raw = p.read_bytes()
    # (4 + 4 + 4) + (4 + 8*19*N + 4), as measured on reference output
    assert len(raw) == 12 + 8 + f.nbytes
    assert raw[:4] == (4).to_bytes(4, "little")
    istat, g = fortio.read_finit(str(p), f.shape)
    assert istat == 7 and np.array_equal(g, f)
    fortio.write_cntd(str(p), f, istep=1234, istat=1, imovie=2)
    (istep, istat, imovie), g = fortio.read_cntd(str(p), f.shape)
    assert (istep, istat, imovie) == (1234, 1, 2) and np.array_equal(g, f)


def test_large_record_subrecords(tmp_path, monkeypatch):
    monkeypatch.setattr(fortio, "_MAXREC", 64)
    a = np.arange(50, dtype="<f8")
    p = tmp_path / "big"
    fortio.write_records(str(p), [a])
    (rec,) = fortio.read_records(str(p))
    assert np.array_equal(np.frombuffer(rec, "<f8"), a)


def test_driver_is_built():
    assert os.path.exists(DRIVER), "build() compiles the Fortran driver with flang"


def _nml(path, **kw):
    items = ", ".join(f"{k}={v}" for k, v in kw.items())
    path.write_text(f"&channel {items} /\n")


def _run(nml, cwd):
    r = subprocess.run([DRIVER, str(nml)], cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.gpu
def test_driver_injected_start_bitwise(tmp_path):
    lx, ly, lz, nsteps = 16, 12, 8, 60
    for d in ("initflow", "cntdflow", "diag", "stat"):
        (tmp_path / d).mkdir()
    ref = O.OracleChannel(lx, ly, lz)
    ref.ux[:], ref.uy[:], ref.uz[:] = workloads.laminar_parabola_noise(lx, ly, lz, seed=21)
    ref.initpop()
    fortio.write_finit(str(tmp_path / "initflow" / "finit.000"), ref.f)
    _nml(tmp_path / "run.nml", lx=lx, ly=ly, lz=lz, nsteps=nsteps, newinitflow=".false.", ndiag=30, nstat=30,
         dirgenr=f"'{tmp_path}/'")
    out = _run(tmp_path / "run.nml", tmp_path)
    assert "MLUPS" in out
    ref.FORCING()
    ref.macrovar()
    ref.run(nsteps)
    (istep, _, _), f = fortio.read_cntd(str(tmp_path / "cntdflow" / f"endrunflow2D16x8.{nsteps:07d}.000"),
                                        ref.f.shape)
    assert istep == nsteps
    assert np.array_equal(f.view(np.uint64), ref.f.view(np.uint64))
    diag = (tmp_path / "diag" / "diag.dat").read_text().split("\n")
    assert len([l for l in diag if l.strip()]) == 2
    assert (tmp_path / "stat" / "profiles2.dat").exists()


@pytest.mark.gpu
def test_driver_fresh_start_prerelax_bitwise(tmp_path):
    lx, ly, lz, nsteps = 32, 8, 6, 40
    for d in ("initflow", "cntdflow", "diag", "stat"):
        (tmp_path / d).mkdir()
    _nml(tmp_path / "run.nml", lx=lx, ly=ly, lz=lz, nsteps=nsteps, newinitflow=".true.", ndiag=0,
         dirgenr=f"'{tmp_path}/'")
    _run(tmp_path / "run.nml", tmp_path)
    ref = O.OracleChannel(lx, ly, lz)
    ref.initvel()
    ref.FORCING()
    ref.initpop()
    ref.prerelax()
    _, finit = fortio.read_finit(str(tmp_path / "initflow" / "finit.000"), ref.f.shape)
    assert np.array_equal(finit.view(np.uint64), ref.f.view(np.uint64))
    ref.macrovar()
    ref.FORCING()
    ref.macrovar()
    ref.run(nsteps)
    _, f = fortio.read_cntd(str(tmp_path / "cntdflow" / f"endrunflow2D16x8.{nsteps:07d}.000"), ref.f.shape)
    assert np.array_equal(f.view(np.uint64), ref.f.view(np.uint64))


@pytest.mark.gpu
def test_driver_perturbation_forcing_bitwise(tmp_path):
    """npforcing > 0 switches on the perturbation forcing of main.f90:148-153:
    FORCINGP before each collision until npforcing, then FORCING."""
    lx, ly, lz, nsteps, npforcing = 16, 12, 8, 40, 30
    for d in ("initflow", "cntdflow", "diag", "stat"):
        (tmp_path / d).mkdir()
    ref = O.OracleChannel(lx, ly, lz)
    ref.ux[:], ref.uy[:], ref.uz[:] = workloads.laminar_parabola_noise(lx, ly, lz, seed=22)
    ref.initpop()
    fortio.write_finit(str(tmp_path / "initflow" / "finit.000"), ref.f)
    _nml(tmp_path / "run.nml", lx=lx, ly=ly, lz=lz, nsteps=nsteps, npforcing=npforcing, newinitflow=".false.",
         ndiag=0, dirgenr=f"'{tmp_path}/'")
    _run(tmp_path / "run.nml", tmp_path)
    ref.FORCING()
    ref.macrovar()
    ref.run_forcingp(0, nsteps, npforcing)
    _, f = fortio.read_cntd(str(tmp_path / "cntdflow" / f"endrunflow2D16x8.{nsteps:07d}.000"), ref.f.shape)
    assert np.array_equal(f.view(np.uint64), ref.f.view(np.uint64))


@pytest.mark.gpu
def test_driver_batched_steps_between_outputs_bitwise(tmp_path):
    """The driver runs the steps between diag outputs as one lbm_run call
    (blocked pairs plus odd single steps); the result is the oracle's."""
    lx, ly, lz, nsteps = 12, 10, 9, 37
    for d in ("initflow", "cntdflow", "diag", "stat"):
        (tmp_path / d).mkdir()
    ref = O.OracleChannel(lx, ly, lz)
    ref.ux[:], ref.uy[:], ref.uz[:] = workloads.laminar_parabola_noise(lx, ly, lz, seed=23)
    ref.initpop()
    fortio.write_finit(str(tmp_path / "initflow" / "finit.000"), ref.f)
    _nml(tmp_path / "run.nml", lx=lx, ly=ly, lz=lz, nsteps=nsteps, newinitflow=".false.", ndiag=10, nstat=15,
         dirgenr=f"'{tmp_path}/'")
    _run(tmp_path / "run.nml", tmp_path)
    ref.FORCING()
    ref.macrovar()
    ref.run(nsteps)
    _, f = fortio.read_cntd(str(tmp_path / "cntdflow" / f"endrunflow2D16x8.{nsteps:07d}.000"), ref.f.shape)
    assert np.array_equal(f.view(np.uint64), ref.f.view(np.uint64))
    diag = (tmp_path / "diag" / "diag.dat").read_text().split("\n")
    assert len([l for l in diag if l.strip()]) == 3
