"""Synthetic workloads of the vacancy-hop path (configs in the reference's
config.yaml terms, flattened as akmc.params_from_config expects).

plain_256      one 256^3 lattice, Fe-1.34Cu-1.30Mn-0.70Ni at%, 2e-4 vacancies,
               573 K, 1e-6 s sector threshold (the plain FP64 benchmark)
precision_512  one 512^3 lattice, same alloy, 2e-4 vacancies, 563 K, per seed
replicas       independent boxes: edge U{16..96} cells, Cu U[0.1,1.5] at%,
               Ni U[0.2,3.5] at%, Mn U[0.3,1.5] at%, vacancies 2e-4 (>= 1),
               T U[500,700] K, one seed set per replica (splitmix64 of a base)
jobs           serving jobs: edge U{32..320}, composition as `replicas`,
               vacancies U[1e-4,5e-4], T U[500,700] K, target events
               log-uniform in [lo, hi]; jobs_with_itl adds dumbbells at 1e-5
               of the sites to a fraction of them
"""
import math

MASK64 = (1 << 64) - 1
ALLOY_REF = [("Fe", 9666), ("Cu", 134), ("Mn", 130), ("Ni", 70)]  # examples/config.yaml order


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return x, z ^ (z >> 31)


class Stream:
    """A splitmix64 stream: reproducible draws without numpy's generators."""

    def __init__(self, seed):
        self.x = seed & MASK64

    def u64(self):
        self.x, z = splitmix64(self.x)
        return z

    def uniform(self, a, b):
        return a + (b - a) * ((self.u64() >> 11) * 2.0 ** -53)

    def integer(self, a, b):  # inclusive
        return a + (self.u64() * (b - a + 1) >> 64)

    def seed31(self):
        return self.u64() & 0x7FFFFFFF


def _cfg(L, alloy, va, T, st, sector_T=3e-7, physics=1e3):
    return dict(box=(L, L, L), alloy=alloy, va_count=va, temperature=T, attempt_freq=6e12, sector_T=sector_T,
                physics_time=physics, seeds=dict(create_types=st.seed31(), create_vacancy=st.seed31(),
                                                 event_selection=0, time_inc=st.seed31()))


def plain_256():
    return dict(box=(256, 256, 256), alloy=ALLOY_REF, va_count=6711,
                seeds=dict(create_types=100565, create_vacancy=16650, event_selection=563264, time_inc=594689),
                temperature=573.0, attempt_freq=6e12, sector_T=1e-6, physics_time=1.0)


def precision_512(k, L=512):
    """Seed k of the precision workload (k = 0..15); L may be reduced."""
    va = int(round(2e-4 * 2 * L ** 3))
    return dict(box=(L, L, L), alloy=[("Fe", 9666), ("Cu", 134), ("Ni", 70), ("Mn", 130)], va_count=va,
                seeds=dict(create_types=100565 + k, create_vacancy=16650 + k, event_selection=0,
                           time_inc=594689 + k),
                temperature=563.0, attempt_freq=6e12, sector_T=3e-7, physics_time=1e3)


def _alloy(st):
    cu = int(round(st.uniform(0.1, 1.5) * 100))
    ni = int(round(st.uniform(0.2, 3.5) * 100))
    mn = int(round(st.uniform(0.3, 1.5) * 100))
    return [("Fe", 10000 - cu - ni - mn), ("Cu", cu), ("Ni", ni), ("Mn", mn)]


def replicas(n=1024, base=0x5EED0001, lo=16, hi=96):
    out = []
    for i in range(n):
        st = Stream(base ^ (i * 0x9E3779B97F4A7C15 & MASK64))
        L = st.integer(lo, hi)
        alloy = _alloy(st)
        va = max(1, int(round(2e-4 * 2 * L ** 3)))
        T = st.uniform(500.0, 700.0)
        out.append(_cfg(L, alloy, va, T, st))
    return out


def jobs(n=4000, base=0x5EED0002, lo=32, hi=320, ev_lo=1e5, ev_hi=1e7):
    """(config, target_events) per job."""
    out = []
    for i in range(n):
        st = Stream(base ^ (i * 0x9E3779B97F4A7C15 & MASK64))
        L = st.integer(lo, hi)
        alloy = _alloy(st)
        va = max(1, int(round(st.uniform(1e-4, 5e-4) * 2 * L ** 3)))
        T = st.uniform(500.0, 700.0)
        target = int(round(math.exp(st.uniform(math.log(ev_lo), math.log(ev_hi)))))
        out.append((_cfg(L, alloy, va, T, st), target))
    return out


def jobs_with_itl(n=4000, base=0x5EED0002, lo=32, hi=320, ev_lo=1e5, ev_hi=1e7, itl_frac=0.1, itl_density=1e-5):
    """(config, target_events, n_itl) per job: the `jobs` law, and a fraction
    itl_frac of the jobs also carry self-interstitial dumbbells at
    itl_density of the sites (at least one)."""
    out = []
    for i, (cfg, target) in enumerate(jobs(n, base, lo, hi, ev_lo, ev_hi)):
        st = Stream((base ^ 0xD0BB) ^ (i * 0x9E3779B97F4A7C15 & MASK64))
        L = cfg["box"][0]
        n_itl = max(1, int(round(itl_density * 2 * L ** 3))) if st.uniform(0.0, 1.0) < itl_frac else 0
        out.append((cfg, target, n_itl))
    return out


DUMBBELL_TYPES = (0x0101, 0x0102, 0x0104, 0x0108, 0x0202, 0x0204, 0x0208, 0x0404, 0x0408, 0x0808)


def typed_lattice(box, seed, f_itl, solutes=(0.0134, 0.0070, 0.0130), f_v=2e-4):
    """A typed lattice for the dumbbell model: Fe with Cu, Ni and Mn atoms
    (fractions `solutes`), vacancies (f_v) and dumbbells of random kind and
    orientation at a fraction f_itl of the sites.  Returns (types, sites,
    orient) in akmc_itl_rates' layout."""
    import numpy as np
    rng = np.random.default_rng(seed)
    n = 2 * box[0] * box[1] * box[2]
    u = rng.random(n)
    t = np.full(n, 1, dtype=np.uint16)
    c = np.cumsum([solutes[0], solutes[1], solutes[2], f_v])
    t[u < c[3]] = 0
    t[u < c[2]] = 8
    t[u < c[1]] = 4
    t[u < c[0]] = 2
    k = max(1, int(round(f_itl * n)))
    sites = np.sort(rng.choice(n, size=k, replace=False)).astype(np.int64)
    t[sites] = rng.choice(np.array(DUMBBELL_TYPES, dtype=np.uint16), size=k)
    orient = rng.choice(np.array([-6, -5, -4, -3, -2, -1, 1, 2, 3, 4, 5, 6], dtype=np.int8), size=k)
    return t, sites, orient
