"""ctypes wrapper of the oracle libraries (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the timed CPU baseline.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

# LatticeTypes::lat_type values (reference src/lattice/lattice_types.h:20-36).
LAT_TYPES = {"V": 0, "Fe": 1, "Cu": 2, "Ni": 4, "Mn": 8,
             "FeFe": 0x0101, "FeCu": 0x0102, "FeNi": 0x0104, "FeMn": 0x0108,
             "CuCu": 0x0202, "CuNi": 0x0204, "CuMn": 0x0208,
             "NiNi": 0x0404, "NiMn": 0x0408, "MnMn": 0x0808}


class RefParams(ctypes.Structure):
    _fields_ = [("box", ctypes.c_int64 * 3), ("ghost", ctypes.c_int64),
                ("n_types", ctypes.c_int32), ("types", ctypes.c_int32 * 32),
                ("ratio", ctypes.c_int64 * 32), ("va_count", ctypes.c_int64),
                ("seed_create_types", ctypes.c_int64), ("seed_create_vacancy", ctypes.c_int64),
                ("seed_time_inc", ctypes.c_int64),
                ("temperature", ctypes.c_double), ("attempt_freq", ctypes.c_double),
                ("sector_T", ctypes.c_double), ("physics_time", ctypes.c_double),
                ("mask_2nn", ctypes.c_uint32)]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load(name):
    path = os.path.join(_BUILD, name)
    if not os.path.exists(path):
        build()
    return ctypes.CDLL(path)


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = _load("librefcpu.so")
        vp = ctypes.c_void_p
        L.ref_create.restype = vp
        L.ref_create.argtypes = [ctypes.POINTER(RefParams)]
        L.ref_destroy.argtypes = [vp]
        L.ref_create_random.argtypes = [vp]
        L.ref_n_sites.restype = ctypes.c_int64
        L.ref_n_sites.argtypes = [vp]
        L.ref_get_sites.argtypes = [vp, ctypes.c_void_p]
        L.ref_set_sites.argtypes = [vp, ctypes.c_void_p]
        L.ref_run.restype = ctypes.c_int
        L.ref_run.argtypes = [vp, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.ref_state.argtypes = [vp, ctypes.c_void_p] + [ctypes.c_void_p] * 6
        L.ref_sector_rates.restype = ctypes.c_int64
        L.ref_sector_rates.argtypes = [vp, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]
        L.ref_delta_e.restype = ctypes.c_double
        L.ref_delta_e.argtypes = [vp, ctypes.c_int64, ctypes.c_int]
        L.ref_rate.restype = ctypes.c_double
        L.ref_rate.argtypes = [vp, ctypes.c_int64, ctypes.c_int]
        L.ref_count.restype = ctypes.c_double
        L.ref_count.argtypes = [vp, ctypes.c_int64, ctypes.c_int]
        L.ref_count_exchange.restype = ctypes.c_double
        L.ref_count_exchange.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.ref_random_atoms_type.argtypes = [vp, ctypes.c_int64]
        L.ref_make_bond.argtypes = [ctypes.c_int, ctypes.c_int]
        L.ref_bond1.restype = ctypes.c_double
        L.ref_bond1.argtypes = [ctypes.c_int]
        L.ref_bond2.restype = ctypes.c_double
        L.ref_bond2.argtypes = [ctypes.c_int]
        L.ref_exp_array.argtypes = [vp, vp, ctypes.c_int64]
        L.ref_exp_array.restype = None
        L.ref_streams.argtypes = [ctypes.c_int64, ctypes.c_int64, vp, vp]
        L.ref_streams.restype = None
        L.ref_set_ghost_images.argtypes = [vp, ctypes.c_int]
        L.ref_set_ghost_images.restype = None
        L.ref_set_list_sweep.argtypes = [vp, ctypes.c_int]
        L.ref_set_list_sweep.restype = None
        L.ref_set_ghost_sync.argtypes = [vp, ctypes.c_int]
        L.ref_set_ghost_sync.restype = None
        L.ref_get_ext_sites.argtypes = [vp, vp, ctypes.c_int64]
        L.ref_get_ext_sites.restype = ctypes.c_int64
        L.ref_region.argtypes = [vp, ctypes.c_int, vp, vp]
        L.ref_region.restype = None
        L.ref_save.argtypes = [vp, ctypes.c_char_p]
        L.ref_save.restype = ctypes.c_int
        L.ref_load.argtypes = [vp, ctypes.c_char_p]
        L.ref_load.restype = ctypes.c_int
        _lib = L
    return _lib


def make_params(cfg):
    """cfg: dict with box, alloy (ordered list of (name, ratio)), va_count,
    seeds (create_types, create_vacancy, time_inc), temperature,
    attempt_freq, sector_T, physics_time, optional ghost / mask_2nn."""
    p = RefParams()
    for d in range(3):
        p.box[d] = int(cfg["box"][d])
    p.ghost = int(cfg.get("ghost", 2))
    alloy = list(cfg["alloy"])
    p.n_types = len(alloy)
    for i, (name, ratio) in enumerate(alloy):
        p.types[i] = LAT_TYPES[name]
        p.ratio[i] = int(ratio)
    p.va_count = int(cfg["va_count"])
    p.seed_create_types = int(cfg["seeds"]["create_types"])
    p.seed_create_vacancy = int(cfg["seeds"]["create_vacancy"])
    p.seed_time_inc = int(cfg["seeds"]["time_inc"])
    p.temperature = float(cfg["temperature"])
    p.attempt_freq = float(cfg["attempt_freq"])
    p.sector_T = float(cfg.get("sector_T", 3.0e-7))
    p.physics_time = float(cfg["physics_time"])
    p.mask_2nn = int(cfg.get("mask_2nn", 0x2F))
    return p


class RefCpu:
    """The reference CPU build's event loop (hash containers, full sweeps).

    ghost_images=True runs the single-rank ghost-image form instead of the
    index wrap: ghost sites hold copies of their periodic images, hops into
    them reach the owner after the sector and changed sites refresh their
    copies (src/pack/*, sublattice.inl:161-163)."""

    def __init__(self, cfg, create=True, ghost_images=False, list_sweep=False):
        self.L = lib()
        self._p = make_params(cfg)
        self.h = self.L.ref_create(ctypes.byref(self._p))
        self.n_sites = self.L.ref_n_sites(self.h)
        if list_sweep:
            self.L.ref_set_list_sweep(self.h, 1)
        if ghost_images:
            self.L.ref_set_ghost_images(self.h, 1)
        if create:
            self.L.ref_create_random(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_destroy(self.h)
            self.h = None

    def sites(self):
        out = np.zeros(self.n_sites, dtype=np.uint8)
        self.L.ref_get_sites(self.h, out.ctypes.data)
        return out

    def ext_sites(self):
        """Flags of the ghost-extended grid, shape (size_z, size_y, size_x)."""
        g = self._p.ghost
        shape = (self._p.box[2] + 2 * g, self._p.box[1] + 2 * g, 2 * (self._p.box[0] + 2 * g))
        out = np.zeros(int(np.prod(shape)), dtype=np.uint8)
        n = self.L.ref_get_ext_sites(self.h, out.ctypes.data, out.size)
        assert n == out.size
        return out.reshape(shape)

    def save(self, path):
        """Writes the whole state (text), for resuming a long run."""
        assert self.L.ref_save(self.h, str(path).encode()) == 0

    def load(self, path):
        assert self.L.ref_load(self.h, str(path).encode()) == 0

    def region(self, s):
        """(lo, hi) cell bounds of sector s on the ghost-extended grid (x not doubled)."""
        lo = np.zeros(3, dtype=np.int64)
        hi = np.zeros(3, dtype=np.int64)
        self.L.ref_region(self.h, int(s), lo.ctypes.data, hi.ctypes.data)
        return lo, hi

    def set_sites(self, sites):
        sites = np.ascontiguousarray(sites, dtype=np.uint8)
        assert sites.size == self.n_sites
        self.L.ref_set_sites(self.h, sites.ctypes.data)

    def run(self, max_events, log=False):
        if log:
            lf = np.zeros(max_events, dtype=np.int64)
            lt = np.zeros(max_events, dtype=np.int64)
            rc = self.L.ref_run(self.h, max_events, lf.ctypes.data, lt.ctypes.data)
            n = self.state()["events"]
            return rc, lf, lt
        return self.L.ref_run(self.h, max_events, None, None), None, None

    def state(self):
        evo = np.zeros(8, dtype=np.float64)
        step = ctypes.c_int64()
        sect = ctypes.c_int32()
        ins = ctypes.c_int32()
        thr = ctypes.c_double()
        st = ctypes.c_double()
        ev = ctypes.c_int64()
        self.L.ref_state(self.h, evo.ctypes.data, ctypes.byref(step), ctypes.byref(sect),
                         ctypes.byref(ins), ctypes.byref(thr), ctypes.byref(st), ctypes.byref(ev))
        return {"evolution": evo, "step": step.value, "sect": sect.value, "in_sector": ins.value,
                "thr": thr.value, "sector_time": st.value, "events": ev.value}

    def sector_rates(self, s, cap=1 << 16):
        sites = np.zeros(cap, dtype=np.int64)
        rates = np.zeros((cap, 8), dtype=np.float64)
        tot = ctypes.c_double()
        n = self.L.ref_sector_rates(self.h, s, cap, sites.ctypes.data, rates.ctypes.data, ctypes.byref(tot))
        assert n <= cap
        return sites[:n], rates[:n], tot.value

    def delta_e(self, site, d):
        return self.L.ref_delta_e(self.h, int(site), int(d))

    def rate(self, site, d):
        return self.L.ref_rate(self.h, int(site), int(d))

    def count(self, site, type_name):
        return self.L.ref_count(self.h, int(site), LAT_TYPES[type_name])

    def count_exchange(self, now_site, source_site, atom, src_atom_type):
        return self.L.ref_count_exchange(self.h, int(now_site), int(source_site), LAT_TYPES[atom],
                                         LAT_TYPES[src_atom_type])

    def random_atoms_type(self, hit):
        return self.L.ref_random_atoms_type(self.h, int(hit))

    def site_index(self, x, y, z):
        """Real-site index of doubled-x coordinates (periodic)."""
        X, Y, Z = 2 * self._p.box[0], self._p.box[1], self._p.box[2]
        return (x % X) + X * ((y % Y) + Y * (z % Z))


def bond1(a, b):
    L = lib()
    return L.ref_bond1(L.ref_make_bond(LAT_TYPES[a], LAT_TYPES[b]))


def bond2(a, b):
    L = lib()
    return L.ref_bond2(L.ref_make_bond(LAT_TYPES[a], LAT_TYPES[b]))


_REF_SRC = os.path.join(_HERE, "_ref", "librefsrc.so")


def refsrc():
    """The reference's own lattice_types.cpp / pair_bond.hpp, compiled from
    source (absent where the reference is not mounted at build time)."""
    if not os.path.exists(_REF_SRC):
        return None
    L = ctypes.CDLL(_REF_SRC)
    L.refsrc_random_atoms_type.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint, ctypes.c_int64]
    L.refsrc_make_bond.argtypes = [ctypes.c_int, ctypes.c_int]
    for f in ("refsrc_combine_to_inter", "refsrc_diff"):
        getattr(L, f).argtypes = [ctypes.c_int, ctypes.c_int]
    L.refsrc_type_info.argtypes = [ctypes.c_int]
    return L


def exp_array(x):
    """The C library's exp over an array."""
    L = lib()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    L.ref_exp_array(x.ctypes.data, y.ctypes.data, ctypes.c_int64(x.size))
    return y


def streams(seed_time_inc, n):
    """(u, -log r) of the first n events as the reference time loop draws them."""
    L = lib()
    u = np.zeros(n, dtype=np.float64)
    l = np.zeros(n, dtype=np.float64)
    L.ref_streams(ctypes.c_int64(seed_time_inc), ctypes.c_int64(n), u.ctypes.data, l.ctypes.data)
    return u, l


# ------------------------------------------------------------------ dumbbells
class Itl:
    """The oracle's restatement of the abvi dumbbell model (orientation.cc,
    defect/itl_list.cpp, rate/itl_rates_solver.cpp, rate/bonds/energy_dumbbell.cpp)."""

    def __init__(self):
        L = lib()
        for f, args in (("ref_ori_avail", [ctypes.c_int, ctypes.c_int]),
                        ("ref_ori_jump", [ctypes.c_int, ctypes.c_int]),
                        ("ref_ori_tran_atom", [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
                        ("ref_ori_from_draw", [ctypes.c_uint32]),
                        ("ref_itl_avail", [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
                        ("ref_itl_rates_index", [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
                        ("ref_itl_first_nn", [ctypes.c_int, ctypes.c_int]),
                        ("ref_itl_rates_status", [ctypes.c_int, ctypes.c_int]),
                        ("ref_combine_to_inter", [ctypes.c_int, ctypes.c_int]),
                        ("ref_type_diff", [ctypes.c_int, ctypes.c_int])):
            getattr(L, f).argtypes = args
            getattr(L, f).restype = ctypes.c_int
        L.ref_itl_update_rates.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.ref_edumb.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int64] + [ctypes.c_void_p] * 3
        L.ref_edumb.restype = ctypes.c_double
        L.ref_count_typed.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]
        L.ref_count_typed.restype = ctypes.c_double
        L.ref_itl_rates.argtypes = ([ctypes.c_void_p] * 2 + [ctypes.c_int64] + [ctypes.c_void_p] * 2 +
                                    [ctypes.c_double, ctypes.c_double] + [ctypes.c_void_p] * 2)
        self.L = L

    def avail(self, o, which=0):
        return self.L.ref_ori_avail(o, which)

    def jump(self, o, b):
        return chr(self.L.ref_ori_jump(o, b))

    def tran_atom(self, o, t, b):
        return self.L.ref_ori_tran_atom(o, t, b)

    def from_draw(self, r):
        return self.L.ref_ori_from_draw(r)

    def itl_avail(self, o, status, nb):
        a = (ctypes.c_int * 8)(*nb)
        return self.L.ref_itl_avail(o, status, a)

    def rates_index(self, b, dirs, up):
        return self.L.ref_itl_rates_index(b, dirs, int(up))

    def first_nn(self, idx, dirs):
        return self.L.ref_itl_first_nn(idx, dirs)

    def rates_status(self, o, avail):
        return self.L.ref_itl_rates_status(o, avail)

    def update_rates(self, o, avail, cb, rates):
        cb = np.ascontiguousarray(cb, dtype=np.float64)
        r = np.ascontiguousarray(rates, dtype=np.float64).copy()
        self.L.ref_itl_update_rates(o, avail, cb.ctypes.data, r.ctypes.data)
        return r

    def combine(self, a, b):
        return self.L.ref_combine_to_inter(a, b)

    def diff(self, t, b):
        return self.L.ref_type_diff(t, b)

    @staticmethod
    def _box(box):
        return np.ascontiguousarray(box, dtype=np.int64)

    def edumb(self, box, types, sites, ori, avail):
        b = self._box(box)
        t = np.ascontiguousarray(types, dtype=np.uint16).copy()
        s = np.ascontiguousarray(sites, dtype=np.int64)
        o = np.ascontiguousarray(ori, dtype=np.int8)
        a = np.ascontiguousarray(avail, dtype=np.uint8)
        return self.L.ref_edumb(b.ctypes.data, t.ctypes.data, s.size, s.ctypes.data, o.ctypes.data, a.ctypes.data)

    def count(self, box, types, site, src):
        b = self._box(box)
        t = np.ascontiguousarray(types, dtype=np.uint16).copy()
        return self.L.ref_count_typed(b.ctypes.data, t.ctypes.data, int(site), int(src))

    def rates(self, box, types, sites, ori, nu, T):
        """(avail[n], rates[n, 8]) of every listed dumbbell."""
        b = self._box(box)
        t = np.ascontiguousarray(types, dtype=np.uint16).copy()
        s = np.ascontiguousarray(sites, dtype=np.int64)
        o = np.ascontiguousarray(ori, dtype=np.int8)
        av = np.zeros(s.size, dtype=np.uint8)
        r = np.zeros((s.size, 8), dtype=np.float64)
        self.L.ref_itl_rates(b.ctypes.data, t.ctypes.data, s.size, s.ctypes.data, o.ctypes.data,
                             float(nu), float(T), av.ctypes.data, r.ctypes.data)
        return av, r


class RefMix:
    """Vacancies and dumbbells on a typed lattice: the abvi model's event loop
    as oracle/refcpu.cpp's refmix restates it (see the comment there)."""

    def __init__(self, box, types, sites, orient, nu=6e12, T=573.0, sector_T=3e-7, physics_time=1.0, ef110=3.53,
                 gen_rate=0.0, seed_time_inc=594689, seed_rec=563264):
        L = lib()
        if not hasattr(L, "_mix_bound"):
            vp, i64, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
            L.refmix_create.restype = vp
            L.refmix_create.argtypes = [vp, dbl, dbl, dbl, dbl, dbl, dbl, i64, i64, vp, i64, vp, vp]
            L.refmix_destroy.argtypes = [vp]
            L.refmix_run.restype = ctypes.c_int
            L.refmix_run.argtypes = [vp, i64, vp, vp, vp, vp]
            L.refmix_state.argtypes = [vp, vp, vp, vp]
            L.refmix_types.argtypes = [vp, vp]
            L.refmix_itl.restype = i64
            L.refmix_itl.argtypes = [vp, i64, vp, vp]
            L.refmix_sector_rates.restype = i64
            L.refmix_sector_rates.argtypes = [vp, ctypes.c_int, i64, vp, vp, vp, vp]
            L.refmix_trans.restype = ctypes.c_int
            L.refmix_trans.argtypes = [ctypes.c_int] * 4
            L._mix_bound = True
        self.L = L
        self.box = tuple(int(v) for v in box)
        self.n_sites = 2 * self.box[0] * self.box[1] * self.box[2]
        b = np.array(self.box, dtype=np.int64)
        t = np.ascontiguousarray(types, dtype=np.uint16)
        s = np.ascontiguousarray(sites, dtype=np.int64)
        o = np.ascontiguousarray(orient, dtype=np.int8)
        assert t.size == self.n_sites and s.size == o.size
        self.h = L.refmix_create(b.ctypes.data, nu, T, sector_T, physics_time, ef110, gen_rate, int(seed_time_inc),
                                 int(seed_rec), t.ctypes.data, s.size, s.ctypes.data, o.ctypes.data)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.refmix_destroy(self.h)
            self.h = None

    def run(self, n, log=False):
        if not log:
            return self.L.refmix_run(self.h, int(n), None, None, None, None), None
        f = np.zeros(n, np.int64)
        t = np.zeros(n, np.int64)
        k = np.zeros(n, np.int32)
        m = ctypes.c_int64()
        rc = self.L.refmix_run(self.h, int(n), f.ctypes.data, t.ctypes.data, k.ctypes.data, ctypes.byref(m))
        return rc, (f[:m.value], t[:m.value], k[:m.value])

    def state(self):
        evo = np.zeros(8)
        ints = np.zeros(7, np.int64)
        dbls = np.zeros(2)
        self.L.refmix_state(self.h, evo.ctypes.data, ints.ctypes.data, dbls.ctypes.data)
        return {"evolution": evo, "step": int(ints[0]), "sect": int(ints[1]), "in_sector": int(ints[2]),
                "events": int(ints[3]), "n_itl": int(ints[4]), "n_vac": int(ints[5]), "recombined": int(ints[6]),
                "thr": float(dbls[0]), "sector_time": float(dbls[1])}

    def types(self):
        out = np.zeros(self.n_sites, np.uint16)
        self.L.refmix_types(self.h, out.ctypes.data)
        return out

    def itl(self):
        n = self.L.refmix_itl(self.h, 0, None, None)
        s = np.zeros(max(n, 1), np.int64)
        o = np.zeros(max(n, 1), np.int8)
        self.L.refmix_itl(self.h, n, s.ctypes.data, o.ctypes.data)
        return s[:n], o[:n]

    def sector_rates(self, sector, cap=1 << 16):
        s = np.zeros(cap, np.int64)
        k = np.zeros(cap, np.int32)
        r = np.zeros((cap, 8))
        tot = ctypes.c_double()
        n = self.L.refmix_sector_rates(self.h, int(sector), cap, s.ctypes.data, k.ctypes.data, r.ctypes.data,
                                       ctypes.byref(tot))
        assert n <= cap
        return s[:n], k[:n], r[:n], tot.value


def mix_trans(o, tag, new_higher, rotate):
    """orientation::trans as the oracle restates it."""
    return lib().refmix_trans(int(o), int(tag), int(bool(new_higher)), int(bool(rotate)))


def mix_place(types, kinds, counts, seed):
    """placeDumbbells restated (oracle): the new types and the placed (site,
    orientation) pairs in placement order."""
    L = lib()
    L.refmix_place.restype = ctypes.c_int64
    L.refmix_place.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                               ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    t = np.array(types, dtype=np.uint16, copy=True)
    k = np.array([LAT_TYPES[x] if isinstance(x, str) else int(x) for x in kinds], dtype=np.int32)
    c = np.array(counts, dtype=np.int64)
    s = np.zeros(int(c.sum()) + 1, np.int64)
    o = np.zeros(int(c.sum()) + 1, np.int8)
    m = L.refmix_place(t.ctypes.data, t.size, k.ctypes.data, c.ctypes.data, len(k), int(seed), s.ctypes.data,
                       o.ctypes.data)
    return t, s[:m], o[:m]
