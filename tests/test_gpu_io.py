"""The reference's Data / Resume HDF5 files on the GPU (include/parties_vof_io.h).

* Layout: the datasets Output_h5_data (IO/Output.c:277-774) and
  Output_h5_resume (:44-276) write for the 2-D VOF_DIFFUSE build, by name,
  type and (NZ, NY, NX) shape, read back with libhdf5 directly (tests/h5read.py).
* Values: each field dataset holds the device field's G range in both k slabs;
  the grid coordinates are Grid.c:845-861's.
* Restart (IO/Resume.c:31-690): a handle resumed from the two files at a step
  boundary continues bit for bit as the uninterrupted run (FFT-PCG pressure,
  cold starts, so no solver state outside the files)."""
import os

import numpy as np
import pytest

import pvof
from h5read import H5File
from helpers import regions

pytestmark = pytest.mark.gpu

DATA_FIELDS = ["/u", "/v", "/w", "/p"] + ["/VOF/" + n for n in (
    "F", "C_L", "C_S", "C_G", "psi", "psi_LG", "lap_C", "bulk_S", "ch_rhs_n", "ch_rhs_nm1", "normal_x",
    "normal_y", "normal_z", "rho", "mu", "f_sigma_old_x", "f_sigma_old_y", "f_sigma_old_z")]
RESUME_FIELDS = ["/Resume/" + n for n in (
    "F", "C_L", "C_S", "ch_rhs_n", "ch_rhs_nm1", "mu", "rho", "flux_x", "flux_y", "flux_z", "normal_x",
    "normal_y", "normal_z", "psi_LG", "f_sigma_old_x", "f_sigma_old_y", "f_sigma_old_z")]


@pytest.fixture(scope="module")
def lib(gpu):
    return pvof.load_lib()


@pytest.fixture(scope="module")
def io(lib):
    return pvof.load_io()


def params():
    return pvof.config1_params(NXM=32, NYM=64, ch_tol=1e-13, CG_ETOL=1e-7, CG_MAXIT=400,
                               pressure_solver=pvof.PV_PSOLVE_FFT_PCG, p_rel_tol=1e-12)


def check(io, rc):
    assert rc == 0, io.pv_io_last_error()


def test_data_and_resume_layout_and_values(lib, io, tmp_path):
    p = params()
    s = pvof.Solver(p, lib)
    s.init_from_F(pvof.ic(lib, "rising_bubble", p))
    s.call("Temporal_int_rk3", 2)
    data, res = str(tmp_path / "Data_1.h5"), str(tmp_path / "Resume_1.h5")
    check(io, io.pv_Output_h5_data(s.h, data.encode()))
    cnt = pvof.IOCounters()
    cnt.noutput, cnt.output_time = 1, 0.25
    for k in range(pvof.PV_IO_NTIMER):
        cnt.timer[k] = 0.5 * k
    check(io, io.pv_Output_h5_resume(s.h, res.encode(), cnt))
    NX, NY, NZ = p.NXM + 1, p.NYM + 1, 2
    st = s.stats()
    f = H5File(data)
    assert f.read("/time")[0] == st["time"]
    assert (f.read("/grid/NX", True)[0], f.read("/grid/NY", True)[0], f.read("/grid/NZ", True)[0]) == (NX, NY, NZ)
    dx = (p.xmax - p.xmin) / p.NXM
    assert np.array_equal(f.read("/grid/xu"), p.xmin + np.arange(NX) * dx)
    assert np.array_equal(f.read("/grid/yc"), p.ymin + (0.5 + np.arange(NY)) * ((p.ymax - p.ymin) / p.NYM))
    for name in DATA_FIELDS:
        assert f.shape(name) == (NZ, NY, NX), name
    G = regions(p)["G"]
    for name, fld in (("/u", "U"), ("/v", "V"), ("/p", "P"), ("/VOF/C_L", "C_L"), ("/VOF/F", "C_L"),
                      ("/VOF/rho", "RHO"), ("/VOF/f_sigma_old_x", "FSX_OLD")):
        a = f.read(name)
        assert np.array_equal(a[0], s.get(fld)[G]), name
        assert np.array_equal(a[1], a[0]), name
    assert not f.read("/w").any() and not f.read("/VOF/C_S").any()
    f.close()
    r = H5File(res)
    assert r.read("/dt")[0] == st["dt"] and r.read("/dt_old")[0] == st["dt_old"]
    assert r.read("/ntime", True)[0] == st["ntime"] and r.read("/noutput", True)[0] == 1
    assert r.shape("/timer") == (34,) and r.read("/timer")[3] == 1.5
    for name in RESUME_FIELDS:
        assert r.shape(name) == (NZ, NY, NX), name
    r.close()
    s.close()


def test_restart_continues_bit_for_bit(lib, io, tmp_path):
    p = params()
    F = pvof.ic(lib, "rising_bubble", p)
    a = pvof.Solver(p, lib)
    a.init_from_F(F)
    a.call("Temporal_int_rk3", 3)
    data, res = str(tmp_path / "Data_3.h5"), str(tmp_path / "Resume.h5")
    check(io, io.pv_Output_h5_data(a.h, data.encode()))
    check(io, io.pv_Output_h5_resume(a.h, res.encode(), None))
    b = pvof.Solver(p, lib)
    cnt = pvof.IOCounters()
    check(io, io.pv_Resume_h5(b.h, res.encode(), data.encode(), cnt))
    sa, sb = a.stats(), b.stats()
    assert (sa["time"], sa["dt"], sa["dt_old"], sa["ntime"]) == (sb["time"], sb["dt"], sb["dt_old"], sb["ntime"])
    a.call("Temporal_int_rk3", 3)
    b.call("Temporal_int_rk3", 3)
    R = regions(p)
    for fld, r in (("C_L", "cell"), ("U", "u"), ("V", "v"), ("P", "cell")):
        assert np.array_equal(a.get(fld)[R[r]], b.get(fld)[R[r]]), fld
    assert a.stats()["time"] == b.stats()["time"]
    a.close()
    b.close()


def test_resume_rejects_wrong_grid(lib, io, tmp_path):
    p = params()
    s = pvof.Solver(p, lib)
    s.init_from_F(pvof.ic(lib, "rising_bubble", p))
    data, res = str(tmp_path / "D.h5"), str(tmp_path / "R.h5")
    check(io, io.pv_Output_h5_data(s.h, data.encode()))
    check(io, io.pv_Output_h5_resume(s.h, res.encode(), None))
    q = pvof.config1_params(NXM=64, NYM=64, pressure_solver=pvof.PV_PSOLVE_FFT_PCG)
    t = pvof.Solver(q, lib)
    assert io.pv_Resume_h5(t.h, res.encode(), data.encode(), None) == -1
    assert b"dimension" in io.pv_io_last_error()
    assert io.pv_Resume_h5(t.h, str(tmp_path / "none.h5").encode(), data.encode(), None) == -1
    s.close()
    t.close()
