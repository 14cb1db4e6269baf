"""The HDF5 Data / Resume module without a GPU: it loads (with the image's
HDF5 1.10), exports its entry points and refuses a null handle."""
import pvof


def test_io_module_loads_and_rejects_null_handle():
    io = pvof.load_io()
    for name in ("pv_Output_h5_data", "pv_Output_h5_resume", "pv_Resume_h5", "pv_io_last_error"):
        assert hasattr(io, name)
    assert io.pv_Output_h5_data(None, b"/nonexistent/x.h5") == -1
    assert b"null handle" in io.pv_io_last_error()
    assert io.pv_Resume_h5(None, b"a.h5", b"b.h5", None) == -1
