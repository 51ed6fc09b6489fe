"""The C-ABI library loads and exports every function include/mdgpu.h declares (no GPU calls). CPU only."""
import ctypes
import subprocess

import mdgpu


def test_library_exports_every_declared_symbol():
    lib = mdgpu.lib()
    declared = mdgpu.header_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing


def test_exported_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", mdgpu.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for s in mdgpu.header_symbols():
        assert s in exported, s


def test_binding_covers_the_header():
    # every declared entry point has a typed ctypes signature in the Python binding
    assert set(mdgpu.header_symbols()) <= set(mdgpu._SIGS)


def test_host_input_generation_matches_oracle():
    """md_fcc_lattice / md_maxwell_boltzmann (host code, no device) reproduce createFCCLattice and
    initializeVelocities bit for bit."""
    import numpy as np

    import oracle as O

    for cells, rho, t, seed in [(5, 0.8442, 0.728, 42), (8, 0.8442, 0.722, 1), (7, 0.75, 1.3, 123)]:
        ref = O.Sim.fcc(cells, rho, t, seed).state()
        a = mdgpu.lattice_constant(rho)
        pos = mdgpu.fcc_lattice(cells, a)
        vel = mdgpu.maxwell_boltzmann(pos.shape[0], t, seed)
        assert np.array_equal(pos, ref["pos"][:, :3])
        assert np.array_equal(vel, ref["vel"])


def test_packed_positions_shapes():
    import numpy as np

    p3 = np.arange(12.0).reshape(4, 3)
    p4 = mdgpu.packed_positions(p3)
    assert p4.shape == (4, 4) and np.array_equal(p4[:, :3], p3) and not p4[:, 3].any()
    q = np.arange(16.0).reshape(4, 4)
    assert np.array_equal(mdgpu.packed_positions(q), q)
    assert mdgpu.packed_positions(np.zeros((0, 3))).shape == (0, 4)


def test_errors_without_touching_the_device():
    lib = mdgpu.lib()
    assert lib.md_state_create(None, 10, 1.0, ctypes.byref(ctypes.c_void_p())) == 1  # MD_ERR_ARG
    assert b"NULL" in lib.md_last_error()
    assert lib.md_list_get_info(None, None) == 1
