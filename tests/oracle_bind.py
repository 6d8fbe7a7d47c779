"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP engine against it, and it is
itself pinned against the reference's regression outputs and the compiled
reference libSteve sources.
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
LIB = ORACLE_DIR / "liboracle.so"
REFLIB = ORACLE_DIR / "_ref" / "libsteve_ref.so"

MS = {"solution": 0, "soln": 0, "fsoln": 0, "front": 1, "back": 2, "up": 3, "down": 4, "bsoln": 5, "all": 6}
RP = {"none": 0, "irrev": 1, "confspread": 2, "bounce": 3, "pgem": 4, "pgemmax": 5, "pgemmaxw": 6,
      "ratio": 7, "unbindrad": 8, "pgem2": 9, "pgemmax2": 10, "ratio2": 11, "offset": 12, "fixed": 13}
# one-letter codes of the configuration language (rxnstring2rp, smolreact.c:79-106)
RP.update({"i": 1, "a": 2, "p": 4, "x": 5, "r": 7, "b": 8, "q": 9, "y": 10, "s": 11, "o": 12, "f": 13})
RNG_SFMT, RNG_COUNTER = 0, 1
PS = {"rect": 0, "tri": 1, "sph": 2, "sphere": 2, "cyl": 3, "cylinder": 3, "hemi": 4, "hemisphere": 4, "disk": 5}
FACE = {"front": 0, "f": 0, "back": 1, "b": 1, "none": 2, "both": 3}
SA = {"reflect": 0, "transmit": 1, "trans": 1, "absorb": 2, "jump": 3, "port": 4, "mult": 5, "no": 6,
      "adsorb": 8, "revdes": 9, "irrevdes": 10, "flip": 11}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        P, D, I, L64 = C.c_void_p, C.c_double, C.c_int, C.c_long
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        sig = {
            "ocmd_timeline": (I, [C.c_char_p, I, dp, ip, C.POINTER(C.c_longlong)]),
            "osim_set_confspread": (I, [P, I, D]),
            "osim_set_bindrad": (I, [P, I, D]),
            "osim_rxnlog_pos": (I, [P, C.POINTER(C.c_double), I]),
            "osim_set_prdpos": (I, [P, I, I, I, dp]),
            "osim_rng_rotation": (None, [C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, dp]),
            "osim_new": (P, [I, dp, dp]),
            "osim_free": (None, [P]),
            "osim_set_wall": (I, [P, I, I, C.c_char]),
            "osim_set_times": (None, [P, D, D, D]),
            "osim_set_seed": (None, [P, L64]),
            "osim_set_rngmode": (None, [P, I, C.c_uint64]),
            "osim_set_partitions": (I, [P, C.c_char_p, D]),
            "osim_set_accuracy": (None, [P, D]),
            "osim_add_species": (I, [P, C.c_char_p]),
            "osim_species_index": (I, [P, C.c_char_p]),
            "osim_set_difc": (None, [P, I, I, D]),
            "osim_add_list": (I, [P, C.c_char_p]),
            "osim_set_list": (I, [P, I, I, I]),
            "osim_add_mols": (I, [P, I, I, dp, dp]),
            "osim_add_mols_at": (I, [P, I, I, dp]),
            "osim_add_reaction": (I, [P, C.c_char_p, I, ip, ip, I, ip, ip, D]),
            "osim_set_revparam": (I, [P, I, I, I, D]),
            "osim_set_rxn_prob": (I, [P, I, I, D]),
            "osim_update": (I, [P]),
            "osim_inject": (I, [P, dp, L64]),
            "osim_step": (I, [P]),
            "osim_run": (I, [P, L64]),
            "osim_time": (D, [P]),
            "osim_nlist": (I, [P]),
            "osim_list_count": (I, [P, I]),
            "osim_get_list": (I, [P, I, ip, ip, dp, C.POINTER(C.c_uint64), ip]),
            "osim_count_species": (I, [P, I]),
            "osim_listmols": (L64, [P, C.c_char_p, L64]),
            "osim_rxn_info": (I, [P, I, I, dp]),
            "osim_box_info": (I, [P, ip, dp, dp]),
            "osim_box_neighbors": (I, [P, I, ip, ip, ip]),
            "osim_rxnlog": (I, [P, C.POINTER(C.c_int64), I]),
            "osim_next_serno": (C.c_uint64, [P]),
            "osim_gausstable": (dp, [P, I]),
            "osim_ngauss": (I, [P]),
            "orxn_numrxnrate": (D, [D, D, D]),
            "orxn_numrxnrateprob": (D, [D, D, D, D]),
            "orxn_actrxnrate": (D, [D, D, D]),
            "orxn_bindingradius": (D, [D, D, D, D, I]),
            "orxn_bindingradiusprob": (D, [D, D, D, D, I, D, dp]),
            "orxn_unbindingradius": (D, [D, D, D, D]),
            "osim_add_surface": (I, [P, C.c_char_p]),
            "osim_set_surface_action": (I, [P, I, I, I, I, I, I]),
            "osim_set_surface_rate": (I, [P, I, I, I, I, I, I, D, I]),
            "osim_add_panel": (I, [P, I, I, C.c_char_p, dp]),
            "osim_set_panel_jump": (I, [P, I, I, I, I, I]),
            "osim_box_panels": (I, [P, I, ip, I]),
            "osurf_geo_spherenormal": (D, [dp, dp, I, I, dp]),
            "osurf_geo_ptinslab": (I, [dp, dp, dp, I]),
            "osurf_geo_linexsphs": (D, [dp, dp, dp, D, I, dp, dp, dp]),
            "osurf_geo_rectxaabb3": (I, [dp, dp, dp, dp, dp, dp]),
            "osurf_geo_sphsxaabb3": (I, [dp, D, dp, dp]),
            "osurf_geo_diskxaabb3": (I, [dp, D, dp, dp, dp]),
            "osurf_geo_linenormal": (D, [dp, dp, dp]),
            "osurf_geo_trinormal": (D, [dp, dp, dp, dp]),
            "osurf_geo_unitcross": (D, [dp, dp, dp, dp, dp]),
            "osurf_geo_ptintriangle2": (I, [dp, dp]),
            "osurf_geo_trixaabb3": (I, [dp, dp, dp, dp, dp, dp]),
            "osurf_geo_linenormal3d": (D, [dp, dp, dp, dp]),
            "osurf_geo_linexcyls": (D, [dp, dp, dp, dp, D, dp, dp, dp]),
            "osurf_geo_linexaabb": (I, [dp, dp, dp, dp, I, I]),
            "osurf_geo_cylsxaabb3": (I, [dp, dp, D, dp, dp]),
            "osurf_geo_hemisxaabb3": (I, [dp, D, dp, dp, dp]),
            "osurf_desorbdist_y": (D, [D, D, I]),
            "osurf_surfaceprob": (D, [D, D, D, D, dp, I]),
            "osurf_transmit": (I, [D, D, D, D, D, dp, dp]),
            "osurf_lookupirrevadsorb": (D, [D]),
            "osurf_lookuprevads": (D, [D, D, dp]),
            "osurf_experfcm1": (D, [D]),
            "omath_randtable": (None, [dp, I]),
            "omath_erfn": (D, [D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class OSim:
    """Thin object wrapper over one oracle simulation."""

    def __init__(self, dim, low, high):
        self.L = lib()
        self.dim = dim
        lo = np.ascontiguousarray(low, dtype=np.float64)
        hi = np.ascontiguousarray(high, dtype=np.float64)
        self.p = self.L.osim_new(dim, _dp(lo), _dp(hi))
        if not self.p:
            raise ValueError("osim_new failed")
        self._keep = []
        self._names = []
        self._lists = []

    def __del__(self):
        if getattr(self, "p", None):
            self.L.osim_free(self.p)
            self.p = None

    def wall(self, d, high, t):
        assert self.L.osim_set_wall(self.p, d, int(high), t.encode()) == 0

    def times(self, t0, t1, dt):
        self.L.osim_set_times(self.p, t0, t1, dt)

    def seed(self, s):
        self.L.osim_set_seed(self.p, s)

    def rngmode(self, mode, key=0):
        self.L.osim_set_rngmode(self.p, mode, key)

    def partitions(self, method, value):
        assert self.L.osim_set_partitions(self.p, method.encode(), value) == 0

    def species(self, name):
        i = self.L.osim_add_species(self.p, name.encode())
        assert i > 0
        self._names.append(name)
        return i

    def index(self, name):
        return self.L.osim_species_index(self.p, name.encode())

    def index_exists(self, i):
        return 0 < i <= len(self._names)

    def list_index(self, name):
        return self._lists.index(name)

    def accuracy(self, a):
        self.L.osim_set_accuracy(self.p, a)

    def difc(self, i, d, ms=0):
        self.L.osim_set_difc(self.p, i, ms, d)

    def add_list(self, name):
        ll = self.L.osim_add_list(self.p, name.encode())
        assert ll >= 0
        self._lists.append(name)
        return ll

    def set_list(self, i, ll, ms=6):
        assert self.L.osim_set_list(self.p, i, ms, ll) == 0

    def add_mols(self, n, i, lo, hi):
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        assert self.L.osim_add_mols(self.p, n, i, _dp(lo), _dp(hi)) == 0

    def add_mols_at(self, i, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, self.dim)
        assert self.L.osim_add_mols_at(self.p, len(pos), i, _dp(pos)) == 0

    def reaction(self, name, rct, prd, rate, rstate=None, pstate=None):
        order = len(rct)
        r = np.array(list(rct) + [0] * (2 - order), dtype=np.int32)
        rs = np.array(rstate if rstate else [0] * order, dtype=np.int32)
        rs = np.concatenate([rs, np.zeros(2 - order, np.int32)])
        p = np.array(list(prd) + [0] * (4 - len(prd)), dtype=np.int32)
        ps = np.array((pstate if pstate else [0] * len(prd)) + [0] * (4 - len(prd)), dtype=np.int32)
        ri = self.L.osim_add_reaction(self.p, name.encode(), order, _ip(r), _ip(rs), len(prd), _ip(p), _ip(ps), rate)
        assert ri >= 0
        return (order, ri)

    def confspread(self, rxn, radius):
        assert rxn[0] == 2 and self.L.osim_set_confspread(self.p, rxn[1], radius) == 0

    def bindrad(self, rxn, radius):
        assert rxn[0] == 2 and self.L.osim_set_bindrad(self.p, rxn[1], radius) == 0

    def prdpos(self, rxn, prd, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        assert self.L.osim_set_prdpos(self.p, rxn[0], rxn[1], prd, _dp(pos)) == 0

    def revparam(self, rxn, kind, value):
        assert self.L.osim_set_revparam(self.p, rxn[0], rxn[1], RP[kind], value) == 0

    def update(self):
        er = self.L.osim_update(self.p)
        if er:
            raise RuntimeError("osim_update error %d" % er)

    def inject(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        self._keep = [g]
        self.L.osim_inject(self.p, _dp(g), g.size // self.dim)

    def step(self):
        return self.L.osim_step(self.p)

    def run(self, maxsteps=-1):
        return self.L.osim_run(self.p, maxsteps)

    def time(self):
        return self.L.osim_time(self.p)

    def nlist(self):
        return self.L.osim_nlist(self.p)

    def get_list(self, ll):
        n = self.L.osim_list_count(self.p, ll)
        ident = np.zeros(n, np.int32)
        ms = np.zeros(n, np.int32)
        pos = np.zeros((n, self.dim), np.float64)
        ser = np.zeros(n, np.uint64)
        box = np.zeros(n, np.int32)
        self.L.osim_get_list(self.p, ll, _ip(ident), _ip(ms), _dp(pos),
                             ser.ctypes.data_as(C.POINTER(C.c_uint64)), _ip(box))
        return dict(ident=ident, mstate=ms, pos=pos, serno=ser, box=box)

    def count(self, i):
        return self.L.osim_count_species(self.p, i)

    def listmols(self):
        n = self.L.osim_listmols(self.p, None, 0)
        buf = C.create_string_buffer(n + 1)
        self.L.osim_listmols(self.p, buf, n + 1)
        return buf.value.decode()

    def rxn_info(self, rxn):
        out = np.zeros(8)
        assert self.L.osim_rxn_info(self.p, rxn[0], rxn[1], _dp(out)) == 0
        return dict(bindrad2=out[0], prob=out[1], unbindrad=out[2], prdpos0=out[3], prdpos1=out[4],
                    rparamt=int(out[5]), rparam=out[6])

    def box_info(self):
        side = np.zeros(3, np.int32)
        bmin = np.zeros(3)
        bsize = np.zeros(3)
        nbox = self.L.osim_box_info(self.p, _ip(side), _dp(bmin), _dp(bsize))
        return nbox, side[:self.dim], bmin[:self.dim], bsize[:self.dim]

    def box_neighbors(self, b):
        nb = np.zeros(27, np.int32)
        wp = np.zeros(27, np.int32)
        mid = np.zeros(1, np.int32)
        n = self.L.osim_box_neighbors(self.p, b, _ip(nb), _ip(wp), _ip(mid))
        return nb[:n], wp[:n], int(mid[0])

    def rxnlog(self):
        n = self.L.osim_rxnlog(self.p, None, 0)
        rows = np.zeros((max(n, 1), 8), np.int64)
        self.L.osim_rxnlog(self.p, rows.ctypes.data_as(C.POINTER(C.c_int64)), n)
        return rows[:n]

    def rxnlog_pos(self):
        n = self.L.osim_rxnlog_pos(self.p, None, 0)
        pos = np.zeros((max(n, 1), self.dim))
        self.L.osim_rxnlog_pos(self.p, _dp(pos), n)
        return pos[:n]

    def add_surface(self, name):
        s = self.L.osim_add_surface(self.p, name.encode())
        assert s >= 0
        return s

    def surface_action(self, s, ident, ms, face, act, newident=-1):
        er = self.L.osim_set_surface_action(self.p, s, ident, ms, face, act, newident)
        assert er == 0, er

    def surface_rate(self, s, ident, ms, ms1, ms2, value, which=1, newident=-5):
        er = self.L.osim_set_surface_rate(self.p, s, ident, ms, ms1, ms2, newident, value, which)
        assert er == 0, er

    def add_panel(self, s, ps, params, axis=None):
        prm = np.zeros(16)
        prm[:len(params)] = params
        p = self.L.osim_add_panel(self.p, s, ps, axis.encode() if axis else None, _dp(prm))
        assert p >= 0, p
        return p

    def panel_jump(self, p1, f1, p2, f2, bidirectional):
        assert self.L.osim_set_panel_jump(self.p, p1, f1, p2, f2, 1 if bidirectional else 0) == 0

    def box_panels(self, b):
        out = np.zeros(256, np.int32)
        n = self.L.osim_box_panels(self.p, b, _ip(out), 256)
        return out[:max(n, 0)].tolist() if n >= 0 else None

    def gausstable(self, flat=True):
        n = self.L.osim_ngauss(self.p)
        ptr = self.L.osim_gausstable(self.p, 1 if flat else 0)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


REFERENCE_TREE = Path("/root/reference/source/libSteve")
_reflib_cache = []


def reflib_available():
    """True where the compiled reference checker can be used: it was built
    and the reference tree it was built from is present (never on a machine
    that only received this repository)."""
    return REFLIB.exists() and REFERENCE_TREE.is_dir()


def reflib():
    """The compiled reference libSteve (checker only), loaded on first use;
    None where it is unavailable."""
    if not reflib_available():
        return None
    if _reflib_cache:
        return _reflib_cache[0]
    L = C.CDLL(str(REFLIB))
    D, I = C.c_double, C.c_int
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    for name, res, args in [
        ("ref_randomize", None, [C.c_long]),
        ("ref_randULI", C.c_ulong, []),
        ("ref_randCOD", D, []),
        ("ref_gausstable", None, [dp, I, I]),
        ("ref_neighbor7", I, [ip, ip, ip, I, ip, ip]),
        ("numrxnrate", D, [D, D, D]),
        ("numrxnrateprob", D, [D, D, D, D]),
        ("actrxnrate", D, [D, D, D]),
        ("bindingradius", D, [D, D, D, D, I]),
        ("bindingradiusprob", D, [D, D, D, D, I, D, dp]),
        ("unbindingradius", D, [D, D, D, D]),
        ("Geo_SphereNormal", D, [dp, dp, I, I, dp]),
        ("Geo_PtInSlab", I, [dp, dp, dp, I]),
        ("Geo_LineXSphs", D, [dp, dp, dp, D, I, dp, dp, dp]),
        ("Geo_RectXaabb3", I, [dp, dp, dp, dp, dp, dp]),
        ("Geo_SphsXaabb3", I, [dp, D, dp, dp]),
        ("Geo_DiskXaabb3", I, [dp, D, dp, dp, dp]),
        ("Geo_LineNormal", D, [dp, dp, dp]),
        ("Geo_TriNormal", D, [dp, dp, dp, dp]),
        ("Geo_UnitCross", D, [dp, dp, dp, dp, dp]),
        ("Geo_PtInTriangle2", I, [C.POINTER(dp), dp]),
        ("Geo_TriXaabb3", I, [dp, dp, dp, dp, dp, dp]),
        ("Geo_LineNormal3D", D, [dp, dp, dp, dp]),
        ("Geo_LineXCyls", D, [dp, dp, dp, dp, D, dp, dp, dp]),
        ("Geo_LineXaabb", I, [dp, dp, dp, dp, I, I]),
        ("Geo_CylsXaabb3", I, [dp, dp, D, dp, dp]),
        ("Geo_HemisXaabb3", I, [dp, D, dp, dp, dp]),
        ("surfaceprob", D, [D, D, D, D, dp, I]),
        ("surfacetransmit", I, [dp, dp, dp, dp, D, D, D]),
        ("lookupirrevadsorb", D, [D, I]),
        ("lookuprevads", D, [D, D, I, dp]),
        ("experfcm1D", D, [D]),
        ("desorbdist", D, [D, I]),
        ("ref_randCCD", D, []),
    ]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _reflib_cache.append(L)
    return L


def command_timeline(script, maxev=200000):
    """The oracle's command scheduler on a script: (times, command positions,
    iterations) of every execution, -1000 - code at the end of each run."""
    L = lib()
    t = np.zeros(maxev)
    c = np.zeros(maxev, np.int32)
    it = np.zeros(maxev, np.int64)
    n = L.ocmd_timeline(script.encode(), maxev, t.ctypes.data_as(C.POINTER(C.c_double)),
                        c.ctypes.data_as(C.POINTER(C.c_int)), it.ctypes.data_as(C.POINTER(C.c_longlong)))
    assert n >= 0, "bad script"
    n = min(n, maxev)
    return t[:n], c[:n], it[:n]
