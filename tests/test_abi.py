"""The C ABI loads and exports every symbol include/traceon_hip.h declares; the Python
bindings match the C struct layouts; host-side logic (mesh, excitation) behaves as the
reference's.  No GPU compute is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "traceon_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"typedef[^;]*;", "", text, flags=re.S)
    text = re.sub(r"struct\s+\w+\s*\{[^}]*\};", "", text, flags=re.S)
    names = set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(\w+)\s*\(", text, flags=re.M))
    return sorted(n for n in names if n not in ("if", "for", "while", "return", "sizeof"))


def declared_data():
    return re.findall(r"extern const \w+ (\w+);", open(HEADER).read())


TABLE = os.path.join(ROOT, "tests", "golden", "backend_functions.json")


def reference_table():
    """The reference's FFI table (traceon/backend/__init__.py:287-328), written by
    tests/golden/make_backend_table.py."""
    import json
    return json.load(open(TABLE))


def test_header_lists_the_reference_ffi():
    names = declared_functions()
    table = reference_table()
    # the table, plus the two traceable field functions the reference passes by name (field.py:816, 1066)
    want = [e["name"] for e in table["functions"]] + ["field_radial_traceable", "field_radial_derivs_traceable"]
    assert len(table["functions"]) == 38
    missing = [n for n in want if n not in names]
    assert not missing, missing
    assert sorted(table["data"]) == sorted(declared_data())


def test_reference_loader_binds_every_table_entry():
    """The reference's loader loop (traceon/backend/__init__.py:341-354) run over this library:
    getattr of every table name, then restype and an argtypes list of the table's length.  It
    raises AttributeError at the first name a library lacks."""
    lib = C.CDLL(os.path.join(ROOT, "traceon_amd", "libtraceon_hip.so"))
    kinds = {"dbl": C.c_double, "sz": C.c_size_t, "bool": C.c_bool, "void": None}
    for e in reference_table()["functions"]:
        f = getattr(lib, e["name"])
        f.restype = kinds[e["result"]]
        f.argtypes = [C.c_void_p] * e["nargs"]
    for d in reference_table()["data"]:
        (C.c_size_t if d == "TRACING_BLOCK_SIZE" else C.c_int).in_dll(lib, d)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(os.path.join(ROOT, "traceon_amd", "libtraceon_hip.so"))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    for d in declared_data():
        assert C.c_int.in_dll(lib, d) is not None
    assert C.c_size_t.in_dll(lib, "TRACING_BLOCK_SIZE").value == 100000
    assert C.c_int.in_dll(lib, "N_QUAD_2D_SYM").value == 16
    assert C.c_int.in_dll(lib, "N_TRIANGLE_QUAD_SYM").value == 12
    assert C.c_int.in_dll(lib, "DERIV_2D_MAX_SYM").value == 9


def test_binding_struct_layouts():
    from traceon_amd import backend
    from traceon_amd.tracing import FieldCompT, FieldProgramT, TraceParamsT
    out = np.zeros(4, dtype=np.uint64)
    backend.backend_lib.tb_struct_sizes(out)
    assert out[0] == C.sizeof(FieldCompT)
    assert out[1] == C.sizeof(FieldProgramT)
    assert out[2] == C.sizeof(TraceParamsT)
    assert out[3] == 48   # RowInfo read as 6 doubles by the solver


def test_mesh_orientation_and_counts():
    from traceon_amd import configs
    from traceon_amd.mesh import rectangle_path_mesh
    m = rectangle_path_mesh(1.0, 2.0, 0.0, 1.0, "r", n_elements=40)
    assert len(m.lines) == 40 and m._normals_outward("r")
    m.ensure_inward_normals("r")
    assert not m._normals_outward("r")
    for fn, n in ((configs.einzel_lens, 3000), (configs.mirror, 20000)):
        mesh, volts = fn(n)
        assert len(mesh.lines) == n
    # element nodes: [start, end, 1/3, 2/3] on straight segments
    v = m.points[m.lines.astype(int)]
    assert np.allclose(v[:, 2], v[:, 0] + (v[:, 1] - v[:, 0]) / 3)
    assert np.allclose(v[:, 3], v[:, 0] + 2 * (v[:, 1] - v[:, 0]) / 3)


def test_excitation_active_elements_and_split():
    from traceon_amd import configs
    from traceon_amd.excitation import Excitation, ExcitationType
    from traceon_amd.mesh import rectangle_surface_mesh
    mesh, volts, diel = configs.small_geometry(120)
    mesh = mesh + rectangle_surface_mesh(3e-3, 4e-3, 0, 1e-3, "coil", 2, 2)
    exc = Excitation(mesh)
    exc.add_voltage(a=-100.0, b=lambda x, y, z: 5.0)
    exc.add_dielectric(d=3.0)
    exc.add_current(coil=2.0)
    verts, names = exc.get_electrostatic_active_elements()
    assert len(verts) == 120 and set(names) == {"a", "b", "d"}
    assert exc.is_electrostatic() and exc.is_magnetostatic()
    elec, mag = exc._split_for_superposition()
    assert set(elec) == {"a", "b"} and set(mag) == {"coil"}
    assert elec["a"].excitation_types["b"][0] == ExcitationType.VOLTAGE_FUN
    assert elec["a"].excitation_types["b"][1](0, 0, 0) == 0.0
    assert mag["coil"].excitation_types["a"][1] == 0.0
    with pytest.raises(AssertionError):
        exc.add_current(a=1.0)   # currents need triangles


def test_product_path_has_no_host_compute():
    """The product package never imports the oracle and computes nothing numeric on the host
    beyond data layout: no numpy.linalg / scipy in the solver, field or tracer modules."""
    import inspect
    from traceon_amd import solver, field, tracing, device, backend
    for mod in (solver, field, tracing, device, backend):
        src = inspect.getsource(mod)
        assert "oracle" not in src.replace("the oracle", "")
        assert not re.search(r"np\.linalg\.(solve|inv|lstsq)\(", src.replace("np.linalg.inv(T)", ""))
        assert not re.search(r"^\s*(import|from)\s+scipy", src, flags=re.M)


def test_clockwise_path_is_flipped_outward():
    """mesher.py:217-219 (ensure_outward_normals on construction): a path given clockwise in
    (r, z) is flipped so that sum(mid_r * n_r * length) > 0, and its 1/3-2/3 nodes follow."""
    from traceon_amd.mesh import path_mesh, circle_mesh
    m = path_mesh([[0.5, 0, -0.25], [0.5, 0, 0.25], [1.0, 0, 0.25]], "tri", n_elements=30, closed=True)
    assert m._normals_outward("tri")
    v = m.points[m.lines.astype(np.int64)]
    # node 2 (1/3) and node 3 (2/3) sit at 1/3 and 2/3 from start to end after the flip
    assert np.allclose(v[:, 2], v[:, 0] + (v[:, 1] - v[:, 0]) / 3) and np.allclose(v[:, 3], v[:, 0] + 2 * (v[:, 1] - v[:, 0]) / 3)
    c = circle_mesh(3.0, 2.0, 1.0, "c", 40)
    assert c._normals_outward("c")
    assert np.allclose(np.linalg.norm(c.points[:, [0, 2]] - [3.0, 2.0], axis=1), 1.0)


def test_library_binds_to_torchs_hip_runtime():
    """Importing the bindings before torch still leaves one HIP and one HSA runtime in the
    process (torch's), the one whose streams and allocations the library is handed."""
    import subprocess
    import sys
    code = ("import traceon_amd.backend\nimport torch\n"
            "libs = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l or 'libhsa-runtime64' in l}\n"
            "print(len(libs)); print(sorted(libs))\n")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    n, libs = out.stdout.splitlines()[:2]
    assert n == "2", libs


def test_off_path_entries_record_an_error():
    """The table's entries that lie off the radial path (kronrod_adaptive, the 3D series) are
    bound by the loader but refuse work: NaN and an error message, no exit (host only)."""
    lib = C.CDLL(os.path.join(ROOT, "traceon_amd", "libtraceon_hip.so"))
    lib.tb_last_error.restype = C.c_char_p
    cb = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)(lambda x, a: x)
    lib.kronrod_adaptive.restype = C.c_double
    lib.kronrod_adaptive.argtypes = [cb.__class__, C.c_double, C.c_double, C.c_void_p, C.c_double, C.c_double]
    lib.tb_clear_error()
    assert np.isnan(lib.kronrod_adaptive(cb, 0.0, 1.0, None, 1e-9, 1e-9))
    assert b"kronrod_adaptive" in lib.tb_last_error()
    lib.potential_3d_derivs.restype = C.c_double
    lib.potential_3d_derivs.argtypes = [C.c_void_p] * 3 + [C.c_size_t]
    lib.tb_clear_error()
    assert np.isnan(lib.potential_3d_derivs(None, None, None, 0))
    assert b"3D" in lib.tb_last_error()


def test_solve_status_maps_to_an_error():
    """A triangular solve whose dependency wait passed its spin limit reports status 1; the
    solver turns it into an exception instead of returning NaN charges (host logic only)."""
    from traceon_amd import device
    device.check_solve_status(0)
    with pytest.raises(RuntimeError, match="spin limit"):
        device.check_solve_status(1)


def test_split_chunk_count_is_defined_for_every_particle_count():
    """tb_trace_split_chunks is host arithmetic: a chunk count in [1, 1024] for any particle
    count, zero included (a trace of no particles launches nothing)."""
    from traceon_amd import backend
    L = backend.backend_lib
    for n in (0, 1, 7, 64, 16384, 65535):
        for ns in (0, 1, 256 * 8, 320000):
            c = L.tb_trace_split_chunks(n, ns)
            assert 1 <= c <= 1024, (n, ns, c)
