"""Test infrastructure: ctypes loader for the CPU oracle (liboracle.so) and for oracle/_ref (the reference
engine's own sources built here, libmdref*.so). Both export the orc_* C ABI of orc_capi.h, so `Sim` drives
either. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {
    "oracle": os.path.join(HERE, "liboracle.so"),
    "ref": os.path.join(HERE, "_ref", "libmdref.so"),
    "ref_fast": os.path.join(HERE, "_ref", "libmdref_fast.so"),
}
_loaded: dict[str, C.CDLL] = {}

_D = C.c_double
_P = C.c_void_p
_DP = C.POINTER(C.c_double)
_U64P = C.POINTER(C.c_uint64)
_IP = C.POINTER(C.c_int)

_SIGS = {
    "orc_last_error": (C.c_char_p, []),
    "orc_set_threads": (None, [C.c_int]),
    "orc_get_threads": (C.c_int, []),
    "orc_create": (_P, [C.c_size_t, _D, _DP, _DP, _DP]),
    "orc_create_fcc": (_P, [C.c_int, _D, _D, C.c_uint32]),
    "orc_create_lattice": (_P, [C.c_int, _D, _D]),
    "orc_clone": (_P, [_P]),
    "orc_destroy": (None, [_P]),
    "orc_size": (C.c_size_t, [_P]),
    "orc_box_length": (_D, [_P]),
    "orc_get_state": (None, [_P, _DP, _DP, _DP, _DP, _U64P]),
    "orc_set_positions": (None, [_P, _DP]),
    "orc_set_velocities": (None, [_P, _DP]),
    "orc_reorder": (C.c_int, [_P, _U64P]),
    "orc_min_image": (None, [_P, _DP, _DP, _DP]),
    "orc_wrap": (_D, [_P, _D]),
    "orc_kinetic_energy": (_D, [_P]),
    "orc_temperature": (_D, [_P]),
    "orc_state_hash": (C.c_uint64, [_P]),
    "orc_cells_build": (C.c_int, [_P, _D, _IP, _IP, _U64P, _U64P]),
    "orc_cells_stencil": (None, [C.c_int, C.c_int, _IP]),
    "orc_forcefield": (None, [_P, _D, _D, _D]),
    "orc_set_mixed": (None, [_P, C.c_int]),
    "orc_set_deterministic": (None, [_P, C.c_int]),
    "orc_energy_shift": (_D, [_P]),
    "orc_forces_allpairs": (C.c_int, [_P, _DP]),
    "orc_forces_list": (C.c_int, [_P, _DP]),
    "orc_virial": (_D, [_P]),
    "orc_inside_pairs": (_D, [_P]),
    "orc_list": (None, [_P, _D, _D, C.c_int]),
    "orc_list_build": (C.c_int, [_P]),
    "orc_list_update": (C.c_int, [_P, _IP]),
    "orc_list_pairs": (C.c_size_t, [_P]),
    "orc_list_rebuilds": (C.c_int, [_P]),
    "orc_list_half": (None, [_P, _U64P, _U64P]),
    "orc_list_reverse": (C.c_int, [_P, _U64P, _U64P]),
    "orc_step_list": (C.c_int, [_P, _D, _DP]),
    "orc_step_allpairs": (C.c_int, [_P, _D, _DP]),
    "orc_rescale": (None, [_P, _D]),
    "orc_bussi": (None, [_P, _D, _D, C.c_uint64]),
    "orc_bussi_apply": (None, [_P, _D]),
    "orc_rdf": (None, [_P, _D, C.c_int]),
    "orc_rdf_sample": (None, [_P]),
    "orc_rdf_values": (None, [_P, _DP, _DP]),
}
# exported by liboracle.so only (no reference function): the fixed-order list and the brute-force pair histogram
_SIGS_ORACLE = {
    "orc_list_fixed": (None, [_P, C.c_int]),
    "orc_pair_histogram": (C.c_int, [_P, _D, C.c_int, _U64P]),
}


def available(kind: str) -> bool:
    return os.path.exists(LIBS[kind])


def lib(kind: str = "oracle") -> C.CDLL:
    if kind not in _loaded:
        path = LIBS[kind]
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
        h = C.CDLL(path)
        for name, (res, args) in list(_SIGS.items()) + (list(_SIGS_ORACLE.items()) if kind == "oracle" else []):
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _loaded[kind] = h
    return _loaded[kind]


def _dp(a):
    return None if a is None else a.ctypes.data_as(_DP)


def _u64p(a):
    return None if a is None else a.ctypes.data_as(_U64P)


class OracleError(RuntimeError):
    pass


class Sim:
    """One System + force field + neighbour list + integrator state, held by the chosen library."""

    def __init__(self, handle, kind: str):
        self.h = handle
        self.kind = kind
        self.L = lib(kind)

    # construction ------------------------------------------------------------------------------------
    @classmethod
    def fcc(cls, cells: int, density: float = 0.8442, temperature: float = 0.728, seed: int = 42,
            kind: str = "oracle") -> "Sim":
        return cls(lib(kind).orc_create_fcc(cells, density, temperature, seed), kind)

    @classmethod
    def lattice(cls, cells: int, a: float, box: float, kind: str = "oracle") -> "Sim":
        return cls(lib(kind).orc_create_lattice(cells, a, box), kind)

    @classmethod
    def from_arrays(cls, pos, box: float, vel=None, mass=None, kind: str = "oracle") -> "Sim":
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 3)
        mass = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64).reshape(n)
        return cls(lib(kind).orc_create(n, box, _dp(pos), _dp(vel), _dp(mass)), kind)

    def clone(self) -> "Sim":
        return Sim(self.L.orc_clone(self.h), self.kind)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.orc_destroy(h)
            self.h = None

    def _check(self, rc: int):
        if rc != 0:
            raise OracleError(self.L.orc_last_error().decode())

    # state ------------------------------------------------------------------------------------------
    @property
    def n(self) -> int:
        return self.L.orc_size(self.h)

    @property
    def box(self) -> float:
        return self.L.orc_box_length(self.h)

    def state(self):
        n = self.n
        pos = np.empty((n, 4))
        vel = np.empty((3, n))
        frc = np.empty((3, n))
        mass = np.empty(n)
        ids = np.empty(n, dtype=np.uint64)
        self.L.orc_get_state(self.h, _dp(pos), _dp(vel), _dp(frc), _dp(mass), _u64p(ids))
        return {"pos": pos, "vel": vel, "force": frc, "mass": mass, "ids": ids}

    def set_positions(self, pos4):
        pos4 = np.ascontiguousarray(pos4, dtype=np.float64).reshape(self.n, 4)
        self.L.orc_set_positions(self.h, _dp(pos4))

    def set_velocities(self, vel):
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(3, self.n)
        self.L.orc_set_velocities(self.h, _dp(vel))

    def reorder(self, order):
        order = np.ascontiguousarray(order, dtype=np.uint64)
        self._check(self.L.orc_reorder(self.h, _u64p(order)))

    def min_image(self, ri, rj):
        ri = np.ascontiguousarray(ri, dtype=np.float64)
        rj = np.ascontiguousarray(rj, dtype=np.float64)
        out = np.empty(3)
        self.L.orc_min_image(self.h, _dp(ri), _dp(rj), _dp(out))
        return out

    def wrap(self, x: float) -> float:
        return self.L.orc_wrap(self.h, x)

    def kinetic_energy(self) -> float:
        return self.L.orc_kinetic_energy(self.h)

    def temperature(self) -> float:
        return self.L.orc_temperature(self.h)

    def state_hash(self) -> int:
        return int(self.L.orc_state_hash(self.h))

    # cells ------------------------------------------------------------------------------------------
    def cells(self, min_cell: float):
        n = self.n
        side = C.c_int(0)
        self._check(self.L.orc_cells_build(self.h, min_cell, C.byref(side), None, None, None))
        s = side.value
        cell_of = np.empty(n, dtype=np.int32)
        start = np.empty(s ** 3 + 1, dtype=np.uint64)
        order = np.empty(n, dtype=np.uint64)
        self._check(self.L.orc_cells_build(self.h, min_cell, C.byref(side), cell_of.ctypes.data_as(_IP),
                                           _u64p(start), _u64p(order)))
        return s, cell_of, start, order

    def stencil(self, side: int, cell: int):
        out = np.empty(27, dtype=np.int32)
        self.L.orc_cells_stencil(side, cell, out.ctypes.data_as(_IP))
        return out

    # force field ------------------------------------------------------------------------------------
    def forcefield(self, eps=1.0, sigma=1.0, rc=2.5, mixed=False, deterministic=False):
        self.L.orc_forcefield(self.h, eps, sigma, rc)
        self.L.orc_set_mixed(self.h, int(mixed))
        self.L.orc_set_deterministic(self.h, int(deterministic))

    def set_mixed(self, on: bool):
        self.L.orc_set_mixed(self.h, int(on))

    def set_deterministic(self, on: bool):
        self.L.orc_set_deterministic(self.h, int(on))

    def forces_allpairs(self) -> float:
        u = C.c_double(0)
        self._check(self.L.orc_forces_allpairs(self.h, C.byref(u)))
        return u.value

    def forces_list(self) -> float:
        u = C.c_double(0)
        self._check(self.L.orc_forces_list(self.h, C.byref(u)))
        return u.value

    @property
    def virial(self) -> float:
        return self.L.orc_virial(self.h)

    @property
    def inside_pairs(self) -> float:
        return self.L.orc_inside_pairs(self.h)

    @property
    def energy_shift(self) -> float:
        return self.L.orc_energy_shift(self.h)

    # neighbour list ---------------------------------------------------------------------------------
    def neighbour_list(self, cutoff=2.5, skin=0.2, reverse=False, fixed_order=False):
        """NeighbourList(cutoff, skin, reverse). fixed_order (oracle only): no re-sort, full rows with the j > i
        part then the j < i part, each in increasing j, and list forces summed along them."""
        self.L.orc_list(self.h, cutoff, skin, int(reverse))
        if fixed_order:
            if self.kind != "oracle":
                raise OracleError("the fixed-order list is an oracle-only mode")
            self.L.orc_list_fixed(self.h, 1)

    def list_build(self):
        self._check(self.L.orc_list_build(self.h))

    def list_update(self) -> bool:
        r = C.c_int(0)
        self._check(self.L.orc_list_update(self.h, C.byref(r)))
        return bool(r.value)

    @property
    def pair_count(self) -> int:
        return self.L.orc_list_pairs(self.h)

    @property
    def rebuilds(self) -> int:
        return self.L.orc_list_rebuilds(self.h)

    def half_list(self):
        offs = np.empty(self.n + 1, dtype=np.uint64)
        neigh = np.empty(max(self.pair_count, 1), dtype=np.uint64)
        self.L.orc_list_half(self.h, _u64p(offs), _u64p(neigh))
        return offs, neigh[: self.pair_count]

    def reverse_index(self):
        roffs = np.empty(self.n + 1, dtype=np.uint64)
        rp = np.empty(max(self.pair_count, 1), dtype=np.uint64)
        if self.L.orc_list_reverse(self.h, _u64p(roffs), _u64p(rp)) != 0:
            raise OracleError("no reverse index")
        return roffs, rp[: self.pair_count]

    # dynamics ---------------------------------------------------------------------------------------
    def step(self, dt=0.005) -> float:
        u = C.c_double(0)
        self._check(self.L.orc_step_list(self.h, dt, C.byref(u)))
        return u.value

    def step_allpairs(self, dt=0.005) -> float:
        u = C.c_double(0)
        self._check(self.L.orc_step_allpairs(self.h, dt, C.byref(u)))
        return u.value

    def rescale(self, t: float):
        self.L.orc_rescale(self.h, t)

    def bussi(self, t: float, tau: float, seed: int):
        self.L.orc_bussi(self.h, t, tau, seed)

    def bussi_apply(self, dt: float):
        self.L.orc_bussi_apply(self.h, dt)

    def rdf(self, rmax: float, bins: int):
        self.L.orc_rdf(self.h, rmax, bins)
        self._bins = bins

    def rdf_sample(self):
        self.L.orc_rdf_sample(self.h)

    def pair_histogram(self, rmax: float, bins: int) -> np.ndarray:
        """Counts of every pair i < j with r^2 < rmax^2 in bin size_t(sqrt(r^2) / (rmax / bins)), by brute force
        (oracle only)."""
        if self.kind != "oracle":
            raise OracleError("the brute-force pair histogram is an oracle-only function")
        out = np.zeros(bins, dtype=np.uint64)
        self._check(self.L.orc_pair_histogram(self.h, rmax, bins, _u64p(out)))
        return out

    def rdf_values(self):
        r = np.empty(self._bins)
        g = np.empty(self._bins)
        self.L.orc_rdf_values(self.h, _dp(r), _dp(g))
        return r, g


def set_threads(kind: str, threads: int):
    lib(kind).orc_set_threads(threads)


def make_liquid(cells: int, steps: int = 300, kind: str = "oracle", density=0.8442, temperature=0.728,
                seed=42, skin=0.2, rc=2.5, dt=0.005) -> Sim:
    """A melted liquid: FCC + MB, list build, list forces, `steps` VV steps with a rescale every 10
    (reference tests/TestHelpers.hpp:39-58)."""
    s = Sim.fcc(cells, density, temperature, seed, kind)
    s.forcefield(1.0, 1.0, rc)
    s.neighbour_list(rc, skin)
    s.list_build()
    s.forces_list()
    for k in range(1, steps + 1):
        s.step(dt)
        if k % 10 == 0:
            s.rescale(temperature)
    return s
