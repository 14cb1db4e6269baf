"""The reference-signature layer (include/parties_vof_shim.h) without a GPU:
it loads, exports every routine its header declares under the reference's
own names, and sends calls on unbound reference pointers down the fatal
path (IO/Display.c:910-937's role) before any device work."""
import ctypes as C
import os
import re

import pytest

import pvof

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "parties_vof_shim.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"typedef[^;]*;", "", src)
    return sorted(set(re.findall(r"\b([A-Za-z][A-Za-z0-9_]*)\s*\(", src)) - {"defined"})


@pytest.fixture(scope="module")
def shim():
    return pvof.load_shim()


def test_shim_exports_every_declared_routine(shim):
    names = declared()
    assert "VOF_DIFFUSE_advect_WENO5" in names and "Temporal_int_rk3" in names
    assert len(names) >= 45
    missing = [n for n in names if not hasattr(shim, n)]
    assert not missing, missing


def test_reference_names_keep_the_reference_prototypes():
    """Each routine of the shim header has the parameter list of its reference
    header (Include/Temporal.h, Dtime.h, Communication.h, VOF_DIFFUSE.h,
    VolumeFraction.h, Velocity.h, Pressure.h), stated here as the reference
    declares them."""
    src = open(os.path.join(ROOT, "include", "parties_vof_shim.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    flat = re.sub(r"\s+", " ", src)
    expect = [
        "int Temporal_int_rk3(Cart3d_bag *data_bag, Debug_trace *dtrace);",
        "void Temporal_int_all_the_equations(Cart3d_bag *data_bag, Debug_trace *dtrace);",
        "double Dtime_cfl(Cart3d_bag *data_bag);",
        "void Communication_update_ghost_nodes_flow_variable(double ***data, char component, int pnodes, "
        "Cart3d_bag *data_bag);",
        "void VOF_DIFFUSE_set_boundary_values(double ***f, Cart3d_bag *db);",
        "void VOF_DIFFUSE_compute_laplacian(double ***in, double ***lap, Cart3d_bag *db);",
        "void VOF_DIFFUSE_advect_WENO5(Cart3d_bag *db, double ***rhs_out);",
        "void VOF_DIFFUSE_solve_implicit_biharmonic(Cart3d_bag *db, double ***rhs_explicit);",
        "void Velocity_update_boundaries(double ***data, char component, int type, Cart3d_bag *data_bag);",
        "int Velocity_solve_cg(Velocity *vel, Cart3d_bag *data_bag);",
        "void Pressure_apply_BCs(double ***phi, MAC_grid *grid, Parameters *params);",
        "double Pressure_compute_velocity_divergence(Cart3d_bag *data_bag);",
        "int Pressure_solve_x_period(Pressure *p, MAC_grid *grid, Parameters *params);",
    ]
    for e in expect:
        assert e in flat, e


def test_unbound_pointers_take_the_fatal_path(shim):
    seen = []

    @pvof.FATAL_FN
    def handler(msg, ctx):
        seen.append(msg.decode())

    shim.pvs_set_fatal(handler, None)
    try:
        bag = C.create_string_buffer(8)
        arr = C.create_string_buffer(8)
        shim.VOF_DIFFUSE_advect_WENO5(C.addressof(bag), C.addressof(arr))
        shim.Pressure_apply_BCs(C.addressof(arr), C.addressof(bag), None)
        assert shim.Velocity_solve_cg(None, C.addressof(bag)) == -1
        assert shim.Dtime_cfl(C.addressof(bag)) == -1.0
        assert len(seen) == 4
        assert "VOF_DIFFUSE_advect_WENO5" in seen[0] and "not bound" in seen[0]
        assert "Pressure_apply_BCs" in seen[1]
    finally:
        shim.pvs_set_fatal(pvof.FATAL_FN(), None)


def test_binding_rejects_bad_arguments(shim):
    bag = C.create_string_buffer(8)
    assert shim.pvs_bind(None, None, None, None, None, None, None) == -1
    assert b"null" in shim.pvs_last_error()
    assert shim.pvs_bind_field(C.addressof(bag), C.addressof(bag), 0, 0) == -1     # bag not bound
    assert shim.pvs_push(C.addressof(bag), C.addressof(bag)) == -1
    assert b"not bound" in shim.pvs_last_error()
    assert shim.pvs_handle(C.addressof(bag)) is None
