"""Python binding of libakmc (include/akmc.h) over ctypes.

The binding is thin: every call goes to the C ABI and raises AkmcError on a
non-zero status.  Loading fails loudly when the HIP library has not been
built; there is no fallback implementation.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libakmc.so")

AKMC_OK = 0
AKMC_ERR_INVALID = -1
AKMC_PREC_FP64 = 0
AKMC_PREC_FP32 = 1
AKMC_STOP_TIME = 1
AKMC_STOP_EVENTS = 2
AKMC_STOP_STEPS = 3
MAX_ALLOY = 32

# LatticeTypes::lat_type values (reference src/lattice/lattice_types.h:20-36).
LAT_TYPES = {"V": 0, "Fe": 1, "Cu": 2, "Ni": 4, "Mn": 8,
             "FeFe": 0x0101, "FeCu": 0x0102, "FeNi": 0x0104, "FeMn": 0x0108,
             "CuCu": 0x0202, "CuNi": 0x0204, "CuMn": 0x0208,
             "NiNi": 0x0404, "NiMn": 0x0408, "MnMn": 0x0808}

# Site flag bits (one uint8 per site).
BIT_V, BIT_CU, BIT_MN, BIT_NI = 1, 2, 4, 8


class AkmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("akmc error %d: %s" % (code, msg))
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("box", ctypes.c_int64 * 3), ("lattice_const", ctypes.c_double),
                ("cutoff_radius", ctypes.c_double), ("n_alloy", ctypes.c_int32),
                ("alloy_type", ctypes.c_int32 * MAX_ALLOY), ("alloy_ratio", ctypes.c_int64 * MAX_ALLOY),
                ("va_count", ctypes.c_int64), ("seed_create_types", ctypes.c_int64),
                ("seed_create_vacancy", ctypes.c_int64), ("seed_event_selection", ctypes.c_int64),
                ("seed_time_inc", ctypes.c_int64), ("temperature", ctypes.c_double),
                ("physics_time", ctypes.c_double), ("attempt_freq", ctypes.c_double),
                ("sector_time", ctypes.c_double), ("precision", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("events", ctypes.c_int64), ("step", ctypes.c_int64), ("sector_pos", ctypes.c_int32),
                ("in_sector", ctypes.c_int32), ("evolution", ctypes.c_double * 8),
                ("sector_threshold", ctypes.c_double), ("sector_time", ctypes.c_double),
                ("stop", ctypes.c_int32), ("sum_errors", ctypes.c_int32), ("seconds", ctypes.c_double),
                ("kernel_seconds", ctypes.c_double), ("launches", ctypes.c_int64),
                ("n_eval", ctypes.c_int64), ("n_fold", ctypes.c_int64), ("n_shift", ctypes.c_int64)]

    def as_dict(self):
        return {"events": self.events, "step": self.step, "sector_pos": self.sector_pos,
                "in_sector": self.in_sector, "evolution": np.array(self.evolution[:]),
                "sector_threshold": self.sector_threshold, "sector_time": self.sector_time,
                "stop": self.stop, "seconds": self.seconds, "kernel_seconds": self.kernel_seconds,
                "launches": self.launches, "n_eval": self.n_eval, "n_fold": self.n_fold,
                "n_shift": self.n_shift, "sum_errors": self.sum_errors}


class Output(ctypes.Structure):
    _fields_ = [("dump_interval", ctypes.c_int64), ("logs_interval", ctypes.c_int64),
                ("dump_file_path", ctypes.c_char * 256), ("logs_file", ctypes.c_char * 256)]


_lib = None


def lib():
    """Loads libakmc.so; raises if it is missing (build it with __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libakmc.so not built at %s; run __graft_entry__.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, i32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    L.akmc_last_error.restype = ctypes.c_char_p
    L.akmc_params_default.argtypes = [P(Params)]
    L.akmc_params_from_yaml.argtypes = [ctypes.c_char_p, P(Params)]
    L.akmc_create.argtypes = [P(Params), i32, P(vp)]
    L.akmc_init_random.argtypes = [vp]
    L.akmc_random_sites.argtypes = [P(Params), vp, i64]
    L.akmc_num_sites.restype = i64
    L.akmc_num_sites.argtypes = [vp]
    L.akmc_upload.argtypes = [vp, vp, i64]
    L.akmc_download.argtypes = [vp, vp, i64]
    L.akmc_run.argtypes = [vp, i64, P(Stats)]
    L.akmc_get_stats.argtypes = [vp, P(Stats)]
    L.akmc_sector_rates.argtypes = [vp, i32, i64, vp, vp, P(dbl), P(i64)]
    L.akmc_site_sectors.argtypes = [P(i64), vp, i64, vp]
    L.akmc_set_event_log.argtypes = [vp, i64]
    L.akmc_get_event_log.argtypes = [vp, vp, vp, i64, P(i64)]
    L.akmc_set_rng_chunk.argtypes = [vp, i64]
    L.akmc_set_check_sums.argtypes = [vp, ctypes.c_int]
    L.akmc_set_select_mode.argtypes = [vp, ctypes.c_int]
    L.akmc_run_steps.argtypes = [vp, i64, i64, P(Stats)]
    L.akmc_checkpoint_save.argtypes = [vp, ctypes.c_char_p]
    L.akmc_checkpoint_load.argtypes = [ctypes.c_char_p, i32, P(vp)]
    L.akmc_output_from_yaml.argtypes = [ctypes.c_char_p, P(Output)]
    L.akmc_dump_path.argtypes = [ctypes.c_char_p, i64, ctypes.c_char_p, i64]
    L.akmc_write_dump.argtypes = [vp, vp, i64, ctypes.c_char_p]
    L.akmc_read_dump.argtypes = [ctypes.c_char_p, vp, P(i64), vp, i64]
    L.akmc_rng_streams.argtypes = [i64, i64, i64, vp, vp]
    L.akmc_sync.argtypes = [i32]
    L.akmc_copy_bench.argtypes = [i32, i64, i32, P(dbl)]
    L.akmc_exp_host.argtypes = [vp, vp, i64]
    L.akmc_exp_device.argtypes = [i32, vp, vp, i64]
    L.akmc_itl_rates.argtypes = [vp, vp, i64, vp, vp, dbl, dbl, i32, vp, vp, P(dbl)]
    L.akmc_itl_rates_noalloc.argtypes = [vp, vp, i64, vp, vp, dbl, dbl, i32, vp, vp, vp]
    L.akmc_pipe_sites.argtypes = [P(Params), ctypes.c_char_p, vp, i64, P(i64)]
    L.akmc_pipe_path_from_yaml.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i64]
    L.akmc_destroy.argtypes = [vp]
    L.akmc_destroy.restype = None
    _lib = L
    return L


def rng_streams(seed_time_inc, first, n):
    """(u, l) of events [first, first + n), as akmc_run consumes them."""
    u = np.zeros(n, dtype=np.float64)
    l = np.zeros(n, dtype=np.float64)
    _check(lib().akmc_rng_streams(int(seed_time_inc), int(first), int(n), u.ctypes.data, l.ctypes.data))
    return u, l


def site_sectors(box, sites):
    """Sector id of each real site index, as the event loops assign them."""
    sites = np.ascontiguousarray(sites, dtype=np.int64)
    out = np.zeros(sites.size, dtype=np.int8)
    b = (ctypes.c_int64 * 3)(*[int(v) for v in box])
    _check(lib().akmc_site_sectors(b, sites.ctypes.data, sites.size, out.ctypes.data))
    return out


def sync(device=0):
    _check(lib().akmc_sync(int(device)))


def copy_bench(nbytes, iters=10, device=0):
    """Seconds per copy of nbytes with the library's streaming copy kernel."""
    sec = ctypes.c_double()
    _check(lib().akmc_copy_bench(int(device), int(nbytes), int(iters), ctypes.byref(sec)))
    return sec.value


def exp_host(x):
    """The rate kernel's exp evaluated on the host (same algorithm as the device)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    _check(lib().akmc_exp_host(x.ctypes.data, y.ctypes.data, x.size))
    return y


def exp_device(x, device=0):
    """The rate kernel's exp evaluated by a HIP kernel."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    _check(lib().akmc_exp_device(int(device), x.ctypes.data, y.ctypes.data, x.size))
    return y


def _check(rc):
    if rc != AKMC_OK:
        raise AkmcError(rc, lib().akmc_last_error().decode())


def params_from_config(cfg):
    """Builds akmc_params from a dict (same keys as the reference config.yaml
    sections, flattened): box, alloy [(name, ratio), ...] in document order,
    va_count, seeds{create_types, create_vacancy, event_selection, time_inc},
    temperature, physics_time, attempt_freq, sector_T, lattice_const,
    cutoff_radius, precision."""
    p = Params()
    _check(lib().akmc_params_default(ctypes.byref(p)))
    for d in range(3):
        p.box[d] = int(cfg["box"][d])
    p.lattice_const = float(cfg.get("lattice_const", 2.85532))
    p.cutoff_radius = float(cfg.get("cutoff_radius", 5.6))
    alloy = list(cfg["alloy"])
    p.n_alloy = len(alloy)
    for i, (name, ratio) in enumerate(alloy):
        p.alloy_type[i] = LAT_TYPES[name]
        p.alloy_ratio[i] = int(ratio)
    p.va_count = int(cfg["va_count"])
    seeds = cfg["seeds"]
    p.seed_create_types = int(seeds["create_types"])
    p.seed_create_vacancy = int(seeds["create_vacancy"])
    p.seed_event_selection = int(seeds.get("event_selection", 0))
    p.seed_time_inc = int(seeds["time_inc"])
    p.temperature = float(cfg["temperature"])
    p.physics_time = float(cfg["physics_time"])
    p.attempt_freq = float(cfg["attempt_freq"])
    p.sector_time = float(cfg.get("sector_T", 3.0e-7))
    p.precision = int(cfg.get("precision", AKMC_PREC_FP64))
    return p


def random_sites(params):
    """Initial occupancy of creation::createRandom computed on the host."""
    if isinstance(params, dict):
        params = params_from_config(params)
    n = 2 * params.box[0] * params.box[1] * params.box[2]
    out = np.zeros(n, dtype=np.uint8)
    _check(lib().akmc_random_sites(ctypes.byref(params), out.ctypes.data, n))
    return out


def pipe_sites(params, pipe_file):
    """Initial occupancy of pipe creation (akmc_pipe_sites): the random
    creation's atoms and a vacancy at each "x y z" triple of pipe_file."""
    if isinstance(params, dict):
        params = params_from_config(params)
    n = 2 * params.box[0] * params.box[1] * params.box[2]
    out = np.zeros(n, dtype=np.uint8)
    nv = ctypes.c_int64()
    _check(lib().akmc_pipe_sites(ctypes.byref(params), str(pipe_file).encode(), out.ctypes.data, n,
                                 ctypes.byref(nv)))
    return out


def pipe_path_from_yaml(path):
    """create.pipe.input_box when pipe creation is the config's create option."""
    buf = ctypes.create_string_buffer(4096)
    _check(lib().akmc_pipe_path_from_yaml(str(path).encode(), buf, len(buf)))
    return buf.value.decode()


def engine_from_yaml(path, device=0):
    """An engine set up as the reference's PKMC::onCreate would set it up from
    config.yaml: random creation, or pipe creation from create.pipe.input_box
    (profiles/pkmc.cpp:165-186)."""
    p = params_from_yaml(str(path))
    try:
        pipe = pipe_path_from_yaml(path)
    except AkmcError:
        return Engine(p, device=device)
    eng = Engine(p, device=device, init_random=False)
    eng.upload(pipe_sites(p, pipe))
    return eng


def params_from_yaml(path):
    p = Params()
    _check(lib().akmc_params_from_yaml(path.encode(), ctypes.byref(p)))
    return p


class Engine:
    """One lattice resident on one GPU, driven through the C ABI."""

    def __init__(self, params, device=0, init_random=True):
        L = lib()
        if isinstance(params, dict):
            params = params_from_config(params)
        self._p = params
        h = ctypes.c_void_p()
        _check(L.akmc_create(ctypes.byref(params), int(device), ctypes.byref(h)))
        self.h = h
        self.n_sites = L.akmc_num_sites(self.h)
        if init_random:
            _check(L.akmc_init_random(self.h))

    @classmethod
    def load(cls, path, device=0):
        """An engine restored from a checkpoint (akmc_checkpoint_load)."""
        L = lib()
        h = ctypes.c_void_p()
        _check(L.akmc_checkpoint_load(str(path).encode(), int(device), ctypes.byref(h)))
        self = cls.__new__(cls)
        self._p = None
        self.h = h
        self.n_sites = L.akmc_num_sites(self.h)
        return self

    def save(self, path):
        """Writes a checkpoint: a loaded engine continues this trajectory bit for bit."""
        _check(lib().akmc_checkpoint_save(self.h, str(path).encode()))

    def close(self):
        if getattr(self, "h", None):
            lib().akmc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def upload(self, sites):
        sites = np.ascontiguousarray(sites, dtype=np.uint8)
        _check(lib().akmc_upload(self.h, sites.ctypes.data, sites.size))

    def download(self):
        out = np.zeros(self.n_sites, dtype=np.uint8)
        _check(lib().akmc_download(self.h, out.ctypes.data, out.size))
        return out

    def run(self, max_events):
        st = Stats()
        _check(lib().akmc_run(self.h, int(max_events), ctypes.byref(st)))
        return st.as_dict()

    def run_steps(self, steps, max_events=1 << 62):
        """Runs until `steps` more time steps are complete (stop AKMC_STOP_STEPS),
        the time loop ends, or max_events events."""
        st = Stats()
        _check(lib().akmc_run_steps(self.h, int(steps), int(max_events), ctypes.byref(st)))
        return st.as_dict()

    def stats(self):
        st = Stats()
        _check(lib().akmc_get_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def sector_rates(self, sector, cap=4096):
        sites = np.zeros(cap, dtype=np.int64)
        rates = np.zeros((cap, 8), dtype=np.float64)
        tot = ctypes.c_double()
        n = ctypes.c_int64()
        _check(lib().akmc_sector_rates(self.h, int(sector), cap, sites.ctypes.data, rates.ctypes.data,
                                       ctypes.byref(tot), ctypes.byref(n)))
        m = min(n.value, cap)
        return sites[:m], rates[:m], tot.value

    def set_rng_chunk(self, events):
        _check(lib().akmc_set_rng_chunk(self.h, int(events)))

    def set_check_sums(self, on=True):
        """Self-check of the running sums after every event (a test aid)."""
        _check(lib().akmc_set_check_sums(self.h, int(bool(on))))

    def set_select_mode(self, mode):
        """0: select from the bounded estimate while the tail repair runs
        (default); 1: always wait for the exact sums; 2: estimate on odd
        events only (a test aid: every mode selects the same events)."""
        _check(lib().akmc_set_select_mode(self.h, int(mode)))

    def set_event_log(self, capacity):
        _check(lib().akmc_set_event_log(self.h, int(capacity)))

    def event_log(self, cap):
        f = np.zeros(cap, dtype=np.int64)
        t = np.zeros(cap, dtype=np.int64)
        n = ctypes.c_int64()
        _check(lib().akmc_get_event_log(self.h, f.ctypes.data, t.ctypes.data, cap, ctypes.byref(n)))
        m = min(n.value, cap)
        return f[:m], t[:m]


# ---------------------------------------------------------------- replica engine
AKMC_INIT_DEVICE = 0
AKMC_INIT_REFERENCE = 1
REP_RUNNING, REP_TIME_DONE, REP_TARGET_DONE = 0, 1, 2
JOB_PENDING, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3


class BatchOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("init", ctypes.c_int32), ("error_stats", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("lds_vcap", ctypes.c_int32), ("waves_per_group", ctypes.c_int32),
                ("order", ctypes.c_int32), ("interleave", ctypes.c_int32), ("slice_events", ctypes.c_int64)]


class BatchStats(ctypes.Structure):
    _fields_ = [("events", ctypes.c_int64), ("seconds", ctypes.c_double), ("kernel_seconds", ctypes.c_double),
                ("launches", ctypes.c_int64), ("running", ctypes.c_int32), ("failed", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ReplicaStats(ctypes.Structure):
    _fields_ = [("events", ctypes.c_int64), ("step", ctypes.c_int64), ("status", ctypes.c_int32),
                ("sector_pos", ctypes.c_int32), ("evolution", ctypes.c_double * 8), ("sector_time", ctypes.c_double),
                ("nvac", ctypes.c_int64), ("n_eval", ctypes.c_int64), ("n_rebuild", ctypes.c_int64),
                ("n_jump", ctypes.c_int64), ("err_max", ctypes.c_double), ("err_mean", ctypes.c_double),
                ("err_n", ctypes.c_int64), ("err_tiny", ctypes.c_int64), ("n_skip", ctypes.c_int64),
                ("busy_seconds", ctypes.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["evolution"] = np.array(self.evolution[:])
        return d


class BatchTuning(ctypes.Structure):
    _fields_ = [("lds_vcap", ctypes.c_int32), ("waves_per_group", ctypes.c_int32), ("order", ctypes.c_int32),
                ("interleave", ctypes.c_int32), ("slice_events", ctypes.c_int64),
                ("events_per_second", ctypes.c_double), ("trials", ctypes.c_int32), ("reserved2", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class ServerOptions(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_int64), ("max_live", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("lds_vcap", ctypes.c_int32), ("waves_per_group", ctypes.c_int32), ("slice_events", ctypes.c_int64),
                ("slice_seconds", ctypes.c_double), ("keep_results", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class JobInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("rep_status", ctypes.c_int32), ("events", ctypes.c_int64),
                ("bytes", ctypes.c_int64), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("admitted", ctypes.c_int64), ("done", ctypes.c_int64),
                ("failed", ctypes.c_int64), ("launches", ctypes.c_int64), ("events", ctypes.c_int64),
                ("live", ctypes.c_int64), ("bytes_live", ctypes.c_int64), ("bytes_peak", ctypes.c_int64),
                ("busy_seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _bind_batch(L):
    vp, i64, i32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    L.akmc_batch_options_default.argtypes = [P(BatchOptions)]
    L.akmc_batch_create.argtypes = [vp, i32, P(BatchOptions), i32, P(vp)]
    L.akmc_batch_run.argtypes = [vp, i64, P(BatchStats)]
    L.akmc_batch_run_for.argtypes = [vp, dbl, P(BatchStats)]
    L.akmc_batch_tune.argtypes = [vp, dbl, P(BatchTuning)]
    L.akmc_batch_get_tuning.argtypes = [vp, P(BatchTuning)]
    L.akmc_batch_tune_trials.argtypes = [vp, vp, ctypes.c_int32, P(ctypes.c_int32)]
    L.akmc_batch_size.argtypes = [vp]
    L.akmc_batch_num_sites.argtypes = [vp, i32]
    L.akmc_batch_num_sites.restype = i64
    L.akmc_batch_replica_stats.argtypes = [vp, i32, P(ReplicaStats)]
    L.akmc_batch_download.argtypes = [vp, i32, vp, i64]
    L.akmc_batch_upload.argtypes = [vp, i32, vp, i64]
    L.akmc_batch_sector_rates.argtypes = [vp, i32, i32, i64, vp, vp, P(dbl), P(i64)]
    L.akmc_fp32_rate_table.argtypes = [dbl, dbl, vp, i32, P(i32), P(i32), P(dbl), vp, vp, vp]
    L.akmc_cu_clusters.argtypes = [vp, vp, i64, vp, P(i64), P(i64), P(i64)]
    L.akmc_batch_destroy.argtypes = [vp]
    L.akmc_batch_destroy.restype = None
    L.akmc_server_options_default.argtypes = [P(ServerOptions)]
    L.akmc_server_create.argtypes = [i32, P(ServerOptions), P(vp)]
    L.akmc_server_submit.argtypes = [vp, P(Params), i64, P(i64)]
    L.akmc_server_job_info.argtypes = [vp, i64, P(JobInfo)]
    L.akmc_server_result.argtypes = [vp, i64, vp, i64]
    L.akmc_server_submit_itl.argtypes = [vp, P(Params), i64, i32, dbl, i64, P(i64)]
    L.akmc_server_result_typed.argtypes = [vp, i64, vp, i64]
    L.akmc_server_wait_idle.argtypes = [vp, dbl]
    L.akmc_server_get_stats.argtypes = [vp, P(ServerStats)]
    L.akmc_server_clock.argtypes = [vp]
    L.akmc_server_clock.restype = dbl
    L.akmc_server_destroy.argtypes = [vp]
    L.akmc_server_destroy.restype = None


_batch_bound = False


def _blib():
    global _batch_bound
    L = lib()
    if not _batch_bound:
        _bind_batch(L)
        _batch_bound = True
    return L


def fp32_rate_table(temperature, nu=6e12):
    """The replica engine's FP32 rate table: (tab, nhi, m_min, q, e1q, e2q, e0q2)."""
    tab = np.zeros(256 + 4096, dtype=np.float32)
    e1q = np.zeros(25, dtype=np.float32)
    e2q = np.zeros(25, dtype=np.float32)
    e0q2 = np.zeros(5, dtype=np.float32)
    nhi, m_min, q = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
    _check(_blib().akmc_fp32_rate_table(float(temperature), float(nu), tab.ctypes.data, tab.size, ctypes.byref(nhi),
                                        ctypes.byref(m_min), ctypes.byref(q), e1q.ctypes.data, e2q.ctypes.data,
                                        e0q2.ctypes.data))
    return tab[:256 + nhi.value].copy(), nhi.value, m_min.value, q.value, e1q, e2q, e0q2


def cu_clusters(sites, box, hist_cap=64):
    """Cu clusters (1nn connectivity) of a site array: dict with the size
    histogram, cluster and atom counts, and the isolated-Cu count."""
    sites = np.ascontiguousarray(sites, dtype=np.uint8)
    b = np.array(box, dtype=np.int64)
    hist = np.zeros(hist_cap, dtype=np.int64)
    nc, ncu, niso = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _check(_blib().akmc_cu_clusters(sites.ctypes.data, b.ctypes.data, hist_cap, hist.ctypes.data, ctypes.byref(nc),
                                    ctypes.byref(ncu), ctypes.byref(niso)))
    return {"hist": hist, "clusters": nc.value, "cu": ncu.value, "isolated": niso.value}


def output_from_yaml(path):
    """The `output` section of a config.yaml (ConfigParsing::parseOutput)."""
    o = Output()
    _check(lib().akmc_output_from_yaml(str(path).encode(), ctypes.byref(o)))
    return {"dump_interval": o.dump_interval, "logs_interval": o.logs_interval,
            "dump_file_path": o.dump_file_path.decode(), "logs_file": o.logs_file.decode()}


def dump_path(pattern, step):
    """fmt::format(dump_file_path, step) for "{}" fields."""
    out = ctypes.create_string_buffer(4096)
    _check(lib().akmc_dump_path(str(pattern).encode(), int(step), out, 4096))
    return out.value.decode()


def write_dump(sites, box, ghost, path):
    """Writes a lattice dump in the reference's format (LatticeDump::dump)."""
    sites = np.ascontiguousarray(sites, dtype=np.uint8)
    b = np.array(box, dtype=np.int64)
    if sites.size != 2 * int(b[0]) * int(b[1]) * int(b[2]):
        raise ValueError("sites do not match the box")
    _check(lib().akmc_write_dump(sites.ctypes.data, b.ctypes.data, int(ghost), str(path).encode()))


def read_dump(path):
    """Reads a lattice dump as the reference's XYZ reader: (sites, box, ghost)."""
    b = np.zeros(3, dtype=np.int64)
    g = ctypes.c_int64()
    _check(lib().akmc_read_dump(str(path).encode(), b.ctypes.data, ctypes.byref(g), None, 0))
    sites = np.zeros(2 * int(b[0]) * int(b[1]) * int(b[2]), dtype=np.uint8)
    _check(lib().akmc_read_dump(str(path).encode(), b.ctypes.data, ctypes.byref(g), sites.ctypes.data, sites.size))
    return sites, tuple(int(x) for x in b), g.value


def run_with_hooks(engine, box, ghost, steps, output, log=print, max_events=1 << 62):
    """The reference's step hook (MEventHook::onStepFinished,
    profiles/m_event_hook.cpp:17-70) around `steps` time steps: after every
    step that is a multiple of 100, the isolated-Cu count is logged; after
    every multiple of output["dump_interval"] (when non-zero) the lattice is
    dumped to dump_path(output["dump_file_path"], step).  The engine runs
    between hook points without stopping.  Returns the last run's stats."""
    # onStepFinished(step) runs after time step `step` (counted from 0,
    # sublattice.inl:61,178) completes, i.e. when the engine's step count
    # reaches step + 1.
    interval = int(output.get("dump_interval", 0))
    st = engine.stats()
    end = st["step"] + int(steps)
    while st["step"] < end:
        s = st["step"]
        hook = -(-s // 100) * 100  # next completed-step index that is a multiple of 100
        if interval:
            hook = min(hook, -(-s // interval) * interval)
        target = min(end, hook + 1)
        st = engine.run_steps(target - s, max_events)
        if st["step"] < target:
            break  # the time loop or the event cap ended first
        done = st["step"] - 1
        if done % 100 == 0:
            log(" step %d finished. now single_cu is %d." % (done, cu_clusters(engine.download(), box)["isolated"]))
        if interval and done % interval == 0:
            write_dump(engine.download(), box, ghost, dump_path(output["dump_file_path"], done))
    return st


def batch_options(**kw):
    o = BatchOptions()
    _check(_blib().akmc_batch_options_default(ctypes.byref(o)))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Batch:
    """Independent replicas advanced together by the replica engine."""

    def __init__(self, configs, device=0, **options):
        L = _blib()
        params = [c if isinstance(c, Params) else params_from_config(c) for c in configs]
        arr = (Params * len(params))(*params)
        self.opts = batch_options(**options)
        h = ctypes.c_void_p()
        _check(L.akmc_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(params), ctypes.byref(self.opts),
                                   int(device), ctypes.byref(h)))
        self.h = h
        self.n = len(params)

    def close(self):
        if getattr(self, "h", None):
            _blib().akmc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run(self, events_per_replica):
        st = BatchStats()
        _check(_blib().akmc_batch_run(self.h, int(events_per_replica), ctypes.byref(st)))
        return st.as_dict()

    def run_for(self, seconds):
        st = BatchStats()
        _check(_blib().akmc_batch_run_for(self.h, float(seconds), ctypes.byref(st)))
        return st.as_dict()

    def tune(self, trial_seconds=0.01):
        t = BatchTuning()
        _check(_blib().akmc_batch_tune(self.h, float(trial_seconds), ctypes.byref(t)))
        return t.as_dict()

    def tuning(self):
        t = BatchTuning()
        _check(_blib().akmc_batch_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def tune_trials(self):
        """Every setting the last tune() tried, with its events per device second."""
        n = ctypes.c_int32()
        _check(_blib().akmc_batch_tune_trials(self.h, None, 0, ctypes.byref(n)))
        arr = (BatchTuning * max(1, n.value))()
        _check(_blib().akmc_batch_tune_trials(self.h, ctypes.cast(arr, ctypes.c_void_p), n.value, ctypes.byref(n)))
        return [arr[i].as_dict() for i in range(n.value)]

    def replica_stats(self, i):
        s = ReplicaStats()
        _check(_blib().akmc_batch_replica_stats(self.h, int(i), ctypes.byref(s)))
        return s.as_dict()

    def num_sites(self, i):
        return _blib().akmc_batch_num_sites(self.h, int(i))

    def download(self, i):
        out = np.zeros(self.num_sites(i), dtype=np.uint8)
        _check(_blib().akmc_batch_download(self.h, int(i), out.ctypes.data, out.size))
        return out

    def upload(self, i, sites):
        sites = np.ascontiguousarray(sites, dtype=np.uint8)
        _check(_blib().akmc_batch_upload(self.h, int(i), sites.ctypes.data, sites.size))

    def sector_rates(self, i, sector, cap=1 << 16):
        sites = np.zeros(cap, dtype=np.int64)
        rates = np.zeros((cap, 8), dtype=np.float64)
        tot = ctypes.c_double()
        n = ctypes.c_int64()
        _check(_blib().akmc_batch_sector_rates(self.h, int(i), int(sector), cap, sites.ctypes.data,
                                               rates.ctypes.data, ctypes.byref(tot), ctypes.byref(n)))
        m = min(n.value, cap)
        return sites[:m], rates[:m], tot.value


def server_options(**kw):
    o = ServerOptions()
    _check(_blib().akmc_server_options_default(ctypes.byref(o)))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Server:
    """The serving loop: jobs are admitted by HBM budget and batched as they arrive."""

    def __init__(self, device=0, **options):
        L = _blib()
        self.opts = server_options(**options)
        h = ctypes.c_void_p()
        _check(L.akmc_server_create(int(device), ctypes.byref(self.opts), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _blib().akmc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def submit(self, config, target_events):
        p = config if isinstance(config, Params) else params_from_config(config)
        job = ctypes.c_int64()
        _check(_blib().akmc_server_submit(self.h, ctypes.byref(p), int(target_events), ctypes.byref(job)))
        return job.value

    def submit_itl(self, config, n_itl, target_events, itl_atom=1, ef110=3.53):
        """A job with n_itl self-interstitial dumbbells (akmc_server_submit_itl)."""
        p = config if isinstance(config, Params) else params_from_config(config)
        job = ctypes.c_int64()
        _check(_blib().akmc_server_submit_itl(self.h, ctypes.byref(p), int(n_itl), int(itl_atom), float(ef110),
                                              int(target_events), ctypes.byref(job)))
        return job.value

    def result_typed(self, job, n_sites):
        out = np.zeros(int(n_sites), dtype=np.uint16)
        _check(_blib().akmc_server_result_typed(self.h, int(job), out.ctypes.data, out.size))
        return out

    def info(self, job):
        i = JobInfo()
        _check(_blib().akmc_server_job_info(self.h, int(job), ctypes.byref(i)))
        return i.as_dict()

    def result(self, job, n_sites):
        out = np.zeros(int(n_sites), dtype=np.uint8)
        _check(_blib().akmc_server_result(self.h, int(job), out.ctypes.data, out.size))
        return out

    def wait_idle(self, timeout=0.0):
        _check(_blib().akmc_server_wait_idle(self.h, float(timeout)))

    def stats(self):
        s = ServerStats()
        _check(_blib().akmc_server_get_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def clock(self):
        return _blib().akmc_server_clock(self.h)


# Dumbbell (self-interstitial) types, LatticeTypes values
# (src/lattice/lattice_types.h:20-36).
LAT_V, LAT_FE, LAT_CU, LAT_NI, LAT_MN = 0, 1, 2, 4, 8


def dumbbell(a, b):
    """LatticeTypes value of the dumbbell of atoms a and b (combineToInter)."""
    a, b = min(a, b), max(a, b)
    return (a << 8) | b


def itl_rates(box, types, sites, orient, attempt_freq, temperature, device=0, timing=None):
    """Rates of the abvi dumbbell model (akmc_itl_rates): (avail[n],
    rates[n, 8]) for the dumbbells at `sites` (ascending) with orientations
    `orient`, on a periodic lattice of LatticeTypes values `types`.  A dict
    passed as `timing` receives the kernels' device seconds."""
    b = np.ascontiguousarray(box, dtype=np.int64)
    t = np.ascontiguousarray(types, dtype=np.uint16)
    s = np.ascontiguousarray(sites, dtype=np.int64)
    o = np.ascontiguousarray(orient, dtype=np.int8)
    if t.size != 2 * int(b[0]) * int(b[1]) * int(b[2]):
        raise AkmcError(AKMC_ERR_INVALID, "types holds %d sites, the box %d" % (t.size, 2 * b[0] * b[1] * b[2]))
    if o.size != s.size:
        raise AkmcError(AKMC_ERR_INVALID, "one orientation per dumbbell")
    av = np.zeros(s.size, dtype=np.uint8)
    r = np.zeros((s.size, 8), dtype=np.float64)
    ks = ctypes.c_double()
    _check(lib().akmc_itl_rates(b.ctypes.data, t.ctypes.data, s.size, s.ctypes.data, o.ctypes.data,
                                float(attempt_freq), float(temperature), int(device), av.ctypes.data,
                                r.ctypes.data, ctypes.byref(ks)))
    if timing is not None:
        timing["kernel_seconds"] = ks.value
    return av, r


ITL_NOALLOC_MAX = 16384


def itl_rates_noalloc(box, d_types, n_itl, d_sites, d_orient, attempt_freq, temperature, d_avail, d_rates,
                      device=0, stream=0):
    """akmc_itl_rates_noalloc on caller-owned device buffers, given as device
    addresses (ints); `stream` is a hipStream_t address (0: null stream)."""
    b = np.ascontiguousarray(box, dtype=np.int64)
    _check(lib().akmc_itl_rates_noalloc(b.ctypes.data, ctypes.c_void_p(d_types), int(n_itl), ctypes.c_void_p(d_sites),
                                        ctypes.c_void_p(d_orient), float(attempt_freq), float(temperature),
                                        int(device), ctypes.c_void_p(stream), ctypes.c_void_p(d_avail),
                                        ctypes.c_void_p(d_rates)))


# ---------------------------------------------------------------- vacancies and dumbbells
class MixParams(ctypes.Structure):
    _fields_ = [("box", ctypes.c_int64 * 3), ("temperature", ctypes.c_double), ("attempt_freq", ctypes.c_double),
                ("sector_time", ctypes.c_double), ("physics_time", ctypes.c_double), ("ef110", ctypes.c_double),
                ("defect_gen_rate", ctypes.c_double), ("seed_time_inc", ctypes.c_int64),
                ("seed_recombine", ctypes.c_int64), ("device_rng", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("vac_capacity", ctypes.c_int64), ("itl_capacity", ctypes.c_int64)]


class MixStats(ctypes.Structure):
    _fields_ = [("events", ctypes.c_int64), ("step", ctypes.c_int64), ("sector_pos", ctypes.c_int32),
                ("in_sector", ctypes.c_int32), ("evolution", ctypes.c_double * 8), ("sector_time", ctypes.c_double),
                ("stop", ctypes.c_int32), ("nvac", ctypes.c_int32), ("nitl", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("n_recombined", ctypes.c_int64), ("n_generated", ctypes.c_int64),
                ("n_itl_events", ctypes.c_int64), ("n_eval", ctypes.c_int64), ("seconds", ctypes.c_double),
                ("kernel_seconds", ctypes.c_double), ("launches", ctypes.c_int64),
                ("n_skipped_steps", ctypes.c_int64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("evolution", "reserved")}
        d["evolution"] = np.array(self.evolution[:])
        return d


def _bind_mix(L):
    if getattr(L, "_mix_bound", False):
        return L
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
    P = ctypes.POINTER
    L.akmc_mix_params_default.argtypes = [P(MixParams)]
    L.akmc_mix_create.argtypes = [i32, vp, vp, vp, vp, vp, ctypes.c_int, P(vp)]
    L.akmc_mix_run.argtypes = [vp, i64, P(i64)]
    L.akmc_mix_stats_of.argtypes = [vp, i32, P(MixStats)]
    L.akmc_mix_download.argtypes = [vp, i32, vp, i64, i64, vp, vp, P(i64)]
    L.akmc_mix_sector_rates.argtypes = [vp, i32, ctypes.c_int, i64, vp, vp, vp, P(ctypes.c_double), P(i64)]
    L.akmc_mix_set_event_log.argtypes = [vp, i64]
    L.akmc_mix_get_event_log.argtypes = [vp, i32, vp, vp, vp, i64, P(i64)]
    L.akmc_mix_destroy.argtypes = [vp]
    L.akmc_mix_destroy.restype = None
    L.akmc_mix_random_lattice.argtypes = [P(Params), vp, vp, i32, i64, vp, i64, i64, vp, vp, P(i64)]
    L._mix_bound = True
    return L


def mix_params(**kw):
    p = MixParams()
    _check(_bind_mix(lib()).akmc_mix_params_default(ctypes.byref(p)))
    for k, v in kw.items():
        if k == "box":
            for d in range(3):
                p.box[d] = int(v[d])
        else:
            setattr(p, k, v)
    return p


def mix_random_lattice(config, itl=(("Fe", 0),), seed=1):
    """creation::createRandom's lattice as LatticeTypes values plus dumbbells
    placed as the reference's placeDumbbells does: (types, sites, orient)."""
    L = _bind_mix(lib())
    prm = config if isinstance(config, Params) else params_from_config(config)
    n = 2 * prm.box[0] * prm.box[1] * prm.box[2]
    kinds = np.array([LAT_TYPES[k] for k, _ in itl], dtype=np.int32)
    counts = np.array([int(c) for _, c in itl], dtype=np.int64)
    t = np.zeros(n, np.uint16)
    cap = int(counts.sum()) + 1
    s = np.zeros(cap, np.int64)
    o = np.zeros(cap, np.int8)
    m = ctypes.c_int64()
    _check(L.akmc_mix_random_lattice(ctypes.byref(prm), kinds.ctypes.data, counts.ctypes.data, len(kinds), int(seed),
                                     t.ctypes.data, n, cap, s.ctypes.data, o.ctypes.data, ctypes.byref(m)))
    return t, s[:m.value], o[:m.value]


class Mix:
    """Lattices with vacancies and dumbbells (akmc_mix_*), one wavefront each."""

    def __init__(self, lattices, device=0):
        """lattices: list of (MixParams, types, sites, orient)."""
        L = _bind_mix(lib())
        self.n = len(lattices)
        self._keep = []
        prm = (MixParams * self.n)(*[x[0] for x in lattices])
        tp = (ctypes.c_void_p * self.n)()
        sp = (ctypes.c_void_p * self.n)()
        op = (ctypes.c_void_p * self.n)()
        ni = np.zeros(self.n, np.int64)
        self.n_sites = []
        for i, (p, t, s, o) in enumerate(lattices):
            t = np.ascontiguousarray(t, np.uint16)
            s = np.ascontiguousarray(s, np.int64)
            o = np.ascontiguousarray(o, np.int8)
            self._keep += [t, s, o]
            tp[i], sp[i], op[i] = t.ctypes.data, s.ctypes.data, o.ctypes.data
            ni[i] = s.size
            self.n_sites.append(t.size)
        h = ctypes.c_void_p()
        _check(L.akmc_mix_create(self.n, ctypes.cast(prm, ctypes.c_void_p), ctypes.cast(tp, ctypes.c_void_p),
                                 ni.ctypes.data, ctypes.cast(sp, ctypes.c_void_p), ctypes.cast(op, ctypes.c_void_p),
                                 int(device), ctypes.byref(h)))
        self.h = h
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            _bind_mix(lib()).akmc_mix_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run(self, max_events):
        tot = ctypes.c_int64()
        _check(_bind_mix(lib()).akmc_mix_run(self.h, int(max_events), ctypes.byref(tot)))
        return tot.value

    def stats(self, i=0):
        st = MixStats()
        _check(_bind_mix(lib()).akmc_mix_stats_of(self.h, int(i), ctypes.byref(st)))
        return st.as_dict()

    def download(self, i=0):
        L = _bind_mix(lib())
        t = np.zeros(self.n_sites[i], np.uint16)
        m = ctypes.c_int64()
        _check(L.akmc_mix_download(self.h, int(i), t.ctypes.data, t.size, 0, None, None, ctypes.byref(m)))
        s = np.zeros(max(m.value, 1), np.int64)
        o = np.zeros(max(m.value, 1), np.int8)
        _check(L.akmc_mix_download(self.h, int(i), None, 0, m.value, s.ctypes.data, o.ctypes.data, ctypes.byref(m)))
        return t, s[:m.value], o[:m.value]

    def sector_rates(self, i, sector, cap=1 << 16):
        L = _bind_mix(lib())
        s = np.zeros(cap, np.int64)
        k = np.zeros(cap, np.int32)
        r = np.zeros((cap, 8))
        tot = ctypes.c_double()
        n = ctypes.c_int64()
        _check(L.akmc_mix_sector_rates(self.h, int(i), int(sector), cap, s.ctypes.data, k.ctypes.data, r.ctypes.data,
                                       ctypes.byref(tot), ctypes.byref(n)))
        m = min(n.value, cap)
        return s[:m], k[:m], r[:m], tot.value

    def set_event_log(self, capacity):
        _check(_bind_mix(lib()).akmc_mix_set_event_log(self.h, int(capacity)))

    def event_log(self, i, cap):
        f = np.zeros(cap, np.int64)
        t = np.zeros(cap, np.int64)
        k = np.zeros(cap, np.int32)
        n = ctypes.c_int64()
        _check(_bind_mix(lib()).akmc_mix_get_event_log(self.h, int(i), f.ctypes.data, t.ctypes.data, k.ctypes.data,
                                                       cap, ctypes.byref(n)))
        m = min(n.value, cap)
        return f[:m], t[:m], k[:m]
