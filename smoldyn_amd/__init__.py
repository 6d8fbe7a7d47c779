"""Python access to libsmoldyn_amd.so, the MI355X engine for Smoldyn's
per-timestep particle path behind the libsmoldyn C API.

The shared library is the product; this module only binds its C-ABI with
ctypes (argument meaning and ErrorCode values as in the reference's
source/Smoldyn/libsmoldyn.h).  It never falls back to a CPU path: loading
fails loudly when the library is missing, and the first time step fails with
the HIP runtime's error when no GPU is present.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
# SMOLAMD_LIBRARY selects another build of the engine library (e.g. a tuning
# variant); the default is the in-tree build.
LIBPATH = Path(os.environ.get("SMOLAMD_LIBRARY", str(_HERE / "libsmoldyn_amd.so")))

# enum ErrorCode (libsmoldyn.h:13-26)
ECok, ECnotify, ECwarning, ECnonexist, ECall, ECmissing, ECbounds, ECsyntax, ECerror, ECmemory, ECbug, ECsame, \
    ECwildcard = 0, -1, -2, -3, -4, -5, -6, -7, -8, -9, -10, -11, -12
# enum MolecState
MSsoln, MSfront, MSback, MSup, MSdown, MSbsoln, MSall, MSnone, MSsome = range(9)
# enum RevParam
RP = {"none": 0, "irrev": 1, "confspread": 2, "bounce": 3, "pgem": 4, "pgemmax": 5, "pgemmaxw": 6, "ratio": 7,
      "unbindrad": 8, "pgem2": 9, "pgemmax2": 10, "ratio2": 11, "offset": 12, "fixed": 13}

_lib = None


class SmoldynError(RuntimeError):
    pass


def build():
    """Compile libsmoldyn_amd.so for gfx950 in-tree (hipcc cross-compiles without a GPU)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", str(_HERE / "csrc"), "-j4"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIBPATH.exists():
        raise OSError("libsmoldyn_amd.so is not built (run smoldyn_amd.build())")
    L = C.CDLL(str(LIBPATH))
    P, D, I, L64 = C.c_void_p, C.c_double, C.c_int, C.c_long
    dp, ip, cp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_char_p
    sig = {
        "smolGetVersion": (D, []),
        "smolGetError": (I, [cp, cp, I]),
        "smolClearError": (None, []),
        "smolSetDebugMode": (None, [I]),
        "smolErrorCodeToString": (cp, [I, cp]),
        "smolNewSim": (P, [I, dp, dp]),
        "smolUpdateSim": (I, [P]),
        "smolRunTimeStep": (I, [P]),
        "smolRunSim": (I, [P]),
        "smolRunSimUntil": (I, [P, D]),
        "smolFreeSim": (I, [P]),
        "smolSetSimTimes": (I, [P, D, D, D]),
        "smolSetTimeStart": (I, [P, D]),
        "smolSetTimeStop": (I, [P, D]),
        "smolSetTimeNow": (I, [P, D]),
        "smolSetTimeStep": (I, [P, D]),
        "smolSetRandomSeed": (I, [P, L64]),
        "smolSetPartitions": (I, [P, cp, D]),
        "smolAddSpecies": (I, [P, cp, cp]),
        "smolGetSpeciesIndex": (I, [P, cp]),
        "smolGetSpeciesIndexNT": (I, [P, cp]),
        "smolGetSpeciesName": (None, [P, I, cp]),
        "smolSetSpeciesMobility": (I, [P, cp, I, D, dp, dp]),
        "smolAddMolList": (I, [P, cp]),
        "smolGetMolListIndex": (I, [P, cp]),
        "smolGetMolListIndexNT": (I, [P, cp]),
        "smolSetMolList": (I, [P, cp, I, cp]),
        "smolSetMaxMolecules": (I, [P, I]),
        "smolAddSolutionMolecules": (I, [P, cp, I, dp, dp]),
        "smolGetMoleculeCount": (I, [P, cp, I]),
        "smolSetBoundaryType": (I, [P, I, I, C.c_char]),
        "smolAddSurface": (I, [P, cp]),
        "smolGetSurfaceIndex": (I, [P, cp]),
        "smolGetSurfaceIndexNT": (I, [P, cp]),
        "smolSetSurfaceAction": (I, [P, cp, I, cp, I, I, cp]),
        "smolSetSurfaceRate": (I, [P, cp, cp, I, I, I, D, cp, I]),
        "smolAddPanel": (I, [P, cp, I, cp, cp, dp]),
        "smolGetPanelIndex": (I, [P, cp, ip, cp]),
        "smolGetPanelIndexNT": (I, [P, cp, ip, cp]),
        "smolSetPanelJump": (I, [P, cp, cp, I, cp, I, I]),
        "smolSetSurfaceSimParams": (I, [P, cp, D]),
        "smolAddReaction": (I, [P, cp, cp, I, cp, I, I, C.POINTER(cp), ip, D]),
        "smolGetReactionIndex": (I, [P, ip, cp]),
        "smolGetReactionIndexNT": (I, [P, ip, cp]),
        "smolSetReactionRate": (I, [P, cp, D, I]),
        "smolSetReactionProducts": (I, [P, cp, I, D, cp, dp]),
        "smolPrepareSimFromFile": (P, [cp, cp, cp]),
        "smolLoadSimFromFile": (I, [cp, cp, C.POINTER(P), cp]),
        "smolReadConfigString": (I, [P, cp, cp]),
        "smolSetOutputPath": (I, [P, cp]),
        "smolAddOutputFile": (I, [P, cp, I, I]),
        "smolAddOutputData": (I, [P, cp]),
        "smolOpenOutputFiles": (I, [P, I]),
        "smolAddCommand": (I, [P, C.c_char, D, D, D, D, cp]),
        "smolAddCommandFromString": (I, [P, cp]),
        "smolGetOutputData": (I, [P, cp, ip, ip, C.POINTER(dp), I]),
        "smolRunCommand": (I, [P, cp]),
        "smolamdCommandTimeline": (I, [cp, I, dp, ip, C.POINTER(C.c_longlong)]),
        "smolamdAddMoleculesAt": (I, [P, cp, I, dp]),
        "smolamdInjectDisplacements": (I, [P, dp, L64]),
        "smolamdInjectDisplacementsNoAlloc": (I, [P, dp, L64]),
        "smolamdSetReactionLog": (I, [P, C.c_char_p, C.c_char_p, C.POINTER(C.c_long), I, I]),
        "smolamdSetPrecision": (I, [P, I]),
        "smolamdGetMolecules": (I, [P, I, ip, dp, C.POINTER(C.c_ulonglong), ip, I]),
        "smolamdGetListCount": (I, [P]),
        "smolamdGetReactionLog": (I, [P, C.POINTER(C.c_longlong), I]),
        "smolamdGetReactionParams": (I, [P, I, I, dp]),
        "smolamdGetBoxes": (I, [P, ip, dp, dp]),
        "smolamdRunSteps": (I, [P, L64, I, dp]),
        "smolamdProfileSteps": (I, [P, L64, dp]),
        "smolamdGetDiagnostics": (I, [P, dp]),
        "smolamdEnsembleNew": (P, [C.POINTER(P), I]),
        "smolamdEnsembleFree": (None, [P]),
        "smolamdEnsembleSize": (I, [P]),
        "smolamdEnsembleRunSteps": (I, [P, L64, I, dp]),
        "smolamdEnsembleGetMoleculeCount": (I, [P, I, cp]),
        "smolamdEnsembleGetMolecules": (I, [P, I, I, ip, dp, C.POINTER(C.c_ulonglong), ip, I]),
        "smolamdEnsembleGetDiagnostics": (I, [P, I, dp]),
        "smolamdBindingRadius": (D, [D, D, D, D, I]),
        "smolamdNumRxnRate": (D, [D, D, D]),
        "smolamdUnbindingRadius": (D, [D, D, D, D]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L.smolSetDebugMode(0)
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def last_error():
    fn = C.create_string_buffer(512)
    msg = C.create_string_buffer(512)
    code = lib().smolGetError(fn, msg, 1)
    return code, fn.value.decode(), msg.value.decode()


_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]
_libc.free.restype = None


def command_timeline(script, maxev=200000):
    """Run the library's command scheduler on a script (smolamdCommandTimeline):
    returns (times, command positions, iterations) of every execution, with a
    position of -1000 - code at the end of each run."""
    L = lib()
    t = np.zeros(maxev)
    c = np.zeros(maxev, np.int32)
    it = np.zeros(maxev, np.int64)
    n = L.smolamdCommandTimeline(script.encode(), maxev, _dp(t), c.ctypes.data_as(C.POINTER(C.c_int)),
                                 it.ctypes.data_as(C.POINTER(C.c_longlong)))
    if n < 0:
        raise SmoldynError("smolamdCommandTimeline: bad script")
    n = min(n, maxev)
    return t[:n], c[:n], it[:n]


def _check(code, what):
    if code < ECwarning:
        ecode, fn, msg = last_error()
        raise SmoldynError("%s failed (%d): %s: %s" % (what, code, fn, msg))
    return code


class Simulation:
    """One simulation behind the C-ABI, mirroring the libsmoldyn call sequence."""

    def __init__(self, low, high):
        self.L = lib()
        lo = np.ascontiguousarray(low, np.float64)
        hi = np.ascontiguousarray(high, np.float64)
        self.dim = len(lo)
        self.p = self.L.smolNewSim(self.dim, _dp(lo), _dp(hi))
        if not self.p:
            raise SmoldynError("smolNewSim: %s" % (last_error(),))

    @classmethod
    def from_file(cls, path, update=True):
        """A simulation read from a configuration file: smolPrepareSimFromFile
        (read and update) or smolLoadSimFromFile (read only)."""
        import os.path
        d, f = os.path.split(str(path))
        self = cls.__new__(cls)
        self.L = lib()
        self.dim = 3
        if update:
            self.p = self.L.smolPrepareSimFromFile((d + "/" if d else "").encode(), f.encode(), b"")
            if not self.p:
                raise SmoldynError("smolPrepareSimFromFile: %s" % (last_error(),))
        else:
            ptr = C.c_void_p(None)
            self.p = None
            _check(self.L.smolLoadSimFromFile((d + "/" if d else "").encode(), f.encode(), C.byref(ptr), b""),
                   "smolLoadSimFromFile")
            self.p = ptr.value
        return self

    def read_config(self, statement, parameters=""):
        """One configuration statement (smolReadConfigString)."""
        _check(self.L.smolReadConfigString(self.p, statement.encode(), parameters.encode()), "smolReadConfigString")

    def close(self):
        if getattr(self, "p", None):
            self.L.smolFreeSim(self.p)
            self.p = None

    def __del__(self):
        self.close()

    # ---- surfaces (libsmoldyn.h:123-136)
    FACES = {"front": 0, "back": 1, "both": 3}
    SHAPES = {"rect": 0, "tri": 1, "triangle": 1, "sph": 2, "sphere": 2, "cyl": 3, "cylinder": 3, "hemi": 4,
              "hemisphere": 4, "disk": 5}
    ACTIONS = {"reflect": 0, "transmit": 1, "absorb": 2, "jump": 3, "port": 4, "mult": 5}
    STATES = {"soln": 0, "fsoln": 0, "front": 1, "back": 2, "up": 3, "down": 4, "bsoln": 5, "all": 6}

    def add_surface(self, name):
        _check(self.L.smolAddSurface(self.p, name.encode()), "smolAddSurface")

    def surface_action(self, surface, face, species, action, state="soln"):
        _check(self.L.smolSetSurfaceAction(self.p, surface.encode(), self.FACES[face], species.encode(),
                                           self.STATES[state], self.ACTIONS[action], None), "smolSetSurfaceAction")

    def surface_rate(self, surface, species, state, state1, state2, rate, newspecies=None, internal=False):
        _check(self.L.smolSetSurfaceRate(self.p, surface.encode(), species.encode(), self.STATES[state],
                                         self.STATES[state1], self.STATES[state2], rate,
                                         newspecies.encode() if newspecies else None, 1 if internal else 0),
               "smolSetSurfaceRate")

    def add_panel(self, surface, shape, params, axis=None, name=None):
        prm = np.zeros(16)
        prm[:len(params)] = params
        _check(self.L.smolAddPanel(self.p, surface.encode(), self.SHAPES[shape], name.encode() if name else None,
                                   axis.encode() if axis else None, _dp(prm)), "smolAddPanel")

    def panel_jump(self, surface, panel1, face1, panel2, face2, bidirectional=True):
        _check(self.L.smolSetPanelJump(self.p, surface.encode(), panel1.encode(), self.FACES[face1], panel2.encode(),
                                       self.FACES[face2], 1 if bidirectional else 0), "smolSetPanelJump")

    def boundary(self, d, high, t):
        _check(self.L.smolSetBoundaryType(self.p, d, high, t.encode()), "smolSetBoundaryType")

    def times(self, t0, t1, dt):
        _check(self.L.smolSetSimTimes(self.p, t0, t1, dt), "smolSetSimTimes")

    def seed(self, s):
        _check(self.L.smolSetRandomSeed(self.p, s), "smolSetRandomSeed")

    def precision(self, bits):
        """Record precision: 64 (FP64 positions) or 32 (FP32 offsets from the
        FP64 origin of each molecule's virtual box, FP64 arithmetic)."""
        _check(self.L.smolamdSetPrecision(self.p, bits), "smolamdSetPrecision")

    def partitions(self, method, value):
        _check(self.L.smolSetPartitions(self.p, method.encode(), value), "smolSetPartitions")

    def species(self, name, mollist=None):
        _check(self.L.smolAddSpecies(self.p, name.encode(), mollist.encode() if mollist else None), "smolAddSpecies")
        return self.L.smolGetSpeciesIndexNT(self.p, name.encode())

    def mobility(self, name, difc, state=MSsoln):
        _check(self.L.smolSetSpeciesMobility(self.p, name.encode(), state, difc, None, None), "smolSetSpeciesMobility")

    def add_list(self, name):
        _check(self.L.smolAddMolList(self.p, name.encode()), "smolAddMolList")

    def set_list(self, species, mollist, state=MSall):
        _check(self.L.smolSetMolList(self.p, species.encode(), state, mollist.encode()), "smolSetMolList")

    def max_molecules(self, n):
        _check(self.L.smolSetMaxMolecules(self.p, n), "smolSetMaxMolecules")

    def add_molecules(self, species, n, lo=None, hi=None):
        lo = None if lo is None else np.ascontiguousarray(lo, np.float64)
        hi = None if hi is None else np.ascontiguousarray(hi, np.float64)
        _check(self.L.smolAddSolutionMolecules(self.p, species.encode(), n, None if lo is None else _dp(lo),
                                               None if hi is None else _dp(hi)), "smolAddSolutionMolecules")

    def add_molecules_at(self, species, pos):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        _check(self.L.smolamdAddMoleculesAt(self.p, species.encode(), len(pos), _dp(pos)), "smolamdAddMoleculesAt")

    def reaction(self, name, reactants, products, rate):
        r1 = reactants[0].encode() if len(reactants) > 0 else None
        r2 = reactants[1].encode() if len(reactants) > 1 else None
        arr = (C.c_char_p * max(1, len(products)))(*[p.encode() for p in products])
        st = (C.c_int * max(1, len(products)))(*([MSsoln] * len(products)))
        _check(self.L.smolAddReaction(self.p, name.encode(), r1, MSsoln, r2, MSsoln, len(products), arr, st, rate),
               "smolAddReaction")

    # ---- run-time commands and output (libsmoldyn.h:85-95)
    def set_output_path(self, path):
        _check(self.L.smolSetOutputPath(self.p, str(path).encode()), "smolSetOutputPath")

    def add_output_file(self, name, suffix=-1, append=False):
        _check(self.L.smolAddOutputFile(self.p, name.encode(), suffix, 1 if append else 0), "smolAddOutputFile")

    def add_output_data(self, name):
        _check(self.L.smolAddOutputData(self.p, name.encode()), "smolAddOutputData")

    def open_output_files(self, overwrite=False):
        _check(self.L.smolOpenOutputFiles(self.p, 1 if overwrite else 0), "smolOpenOutputFiles")

    def add_command(self, command, timing, on=0.0, off=0.0, step=0.0, multiplier=0.0):
        _check(self.L.smolAddCommand(self.p, timing.encode()[:1], on, off, step, multiplier, command.encode()),
               "smolAddCommand")

    def add_command_str(self, line):
        _check(self.L.smolAddCommandFromString(self.p, line.encode()), "smolAddCommandFromString")

    def output_data(self, name, erase=True):
        """Rows of data table `name` as an (nrow, ncol) array (smolGetOutputData)."""
        nr, nc = C.c_int(0), C.c_int(0)
        arr = C.POINTER(C.c_double)()
        _check(self.L.smolGetOutputData(self.p, name.encode(), C.byref(nr), C.byref(nc), C.byref(arr),
                                        1 if erase else 0), "smolGetOutputData")
        try:
            n = nr.value * nc.value
            out = np.ctypeslib.as_array(arr, shape=(n,)).copy() if n > 0 else np.zeros(0)
        finally:
            _libc.free(C.cast(arr, C.c_void_p))
        return out.reshape(nr.value, nc.value)

    def run_command(self, command):
        return _check(self.L.smolRunCommand(self.p, command.encode()), "smolRunCommand")

    def product_placement(self, name, kind, param=0.0, product=None, position=None):
        """smolSetReactionProducts; offset and fixed placement take one
        product's position per call."""
        pos = None if position is None else np.ascontiguousarray(position, np.float64)
        _check(self.L.smolSetReactionProducts(self.p, name.encode(), RP[kind], param,
                                              product.encode() if product else None,
                                              None if pos is None else _dp(pos)), "smolSetReactionProducts")

    def binding_radius(self, name, radius):
        """smolSetReactionRate with isinternal 1: the binding radius (for
        confspread reactions, the conformational-spread radius)."""
        _check(self.L.smolSetReactionRate(self.p, name.encode(), radius, 1), "smolSetReactionRate")

    def reaction_index(self, name):
        order = C.c_int(-1)
        r = self.L.smolGetReactionIndexNT(self.p, C.byref(order), name.encode())
        return order.value, r

    def update(self):
        _check(self.L.smolUpdateSim(self.p), "smolUpdateSim")

    def run_time_step(self):
        return _check(self.L.smolRunTimeStep(self.p), "smolRunTimeStep")

    def run(self):
        return _check(self.L.smolRunSim(self.p), "smolRunSim")

    def run_until(self, t):
        return _check(self.L.smolRunSimUntil(self.p, t), "smolRunSimUntil")

    def inject_step(self, g):
        g = np.ascontiguousarray(g, np.float64).reshape(-1, 3)
        return _check(self.L.smolamdInjectDisplacements(self.p, _dp(g), len(g)), "smolamdInjectDisplacements")

    def set_reaction_log(self, filename, reaction="all", serials=None, on=True):
        """reaction_log statement: log the events of `reaction` whose products'
        serial numbers are listed (None: every event) to output file `filename`."""
        if serials is None:
            arr, n = None, -1
        else:
            arr = (C.c_long * max(len(serials), 1))(*serials)
            n = len(serials)
        _check(self.L.smolamdSetReactionLog(self.p, filename.encode(), reaction.encode(), arr, n, 1 if on else 0),
               "smolamdSetReactionLog")

    def inject_step_noalloc(self, g, nmol=None):
        """inject_step through the entry point that allocates no device memory."""
        g = np.ascontiguousarray(g, np.float64).reshape(-1, 3)
        n = len(g) if nmol is None else nmol
        return _check(self.L.smolamdInjectDisplacementsNoAlloc(self.p, _dp(g), n), "smolamdInjectDisplacementsNoAlloc")

    def run_steps(self, n, graph=True):
        ms = C.c_double(0)
        _check(self.L.smolamdRunSteps(self.p, n, 1 if graph else 0, C.byref(ms)), "smolamdRunSteps")
        return ms.value

    def profile_steps(self, n):
        """Per-kernel ms from each kernel's own dispatch start / stop events:
        diffuse, scan, scatter, first molsort (relist), pair screen, exact
        test, resolve, tile binning pass 1, tile offsets, tile pass 2."""
        out = np.zeros(10)
        _check(self.L.smolamdProfileSteps(self.p, n, _dp(out)), "smolamdProfileSteps")
        return out

    def count(self, species=None, state=MSsoln):
        n = self.L.smolGetMoleculeCount(self.p, species.encode() if species else b"all", state)
        _check(n if n < 0 else 0, "smolGetMoleculeCount")
        return n

    def nlists(self):
        return _check(self.L.smolamdGetListCount(self.p), "smolamdGetListCount")

    def molecules(self, ll, maxmol=1 << 26):
        n = self.L.smolamdGetMolecules(self.p, ll, None, None, None, None, 0)
        _check(n if n < 0 else 0, "smolamdGetMolecules")
        ident = np.zeros(n, np.int32)
        pos = np.zeros((n, 3), np.float64)
        ser = np.zeros(n, np.uint64)
        box = np.zeros(n, np.int32)
        self.L.smolamdGetMolecules(self.p, ll, ident.ctypes.data_as(C.POINTER(C.c_int)), _dp(pos),
                                   ser.ctypes.data_as(C.POINTER(C.c_ulonglong)),
                                   box.ctypes.data_as(C.POINTER(C.c_int)), n)
        return dict(ident=ident, pos=pos, serno=ser, box=box)

    def reaction_log(self):
        n = self.L.smolamdGetReactionLog(self.p, None, 0)
        rows = np.zeros((max(n, 1), 6), np.int64)
        self.L.smolamdGetReactionLog(self.p, rows.ctypes.data_as(C.POINTER(C.c_longlong)), n)
        return rows[:n]

    def reaction_params(self, order, r):
        out = np.zeros(6)
        _check(self.L.smolamdGetReactionParams(self.p, order, r, _dp(out)), "smolamdGetReactionParams")
        return dict(bindrad2=out[0], prob=out[1], unbindrad=out[2], prdpos0=out[3], prdpos1=out[4],
                    rparamt=int(out[5]))

    def boxes(self):
        side = np.zeros(3, np.int32)
        bmin = np.zeros(3)
        bsize = np.zeros(3)
        n = self.L.smolamdGetBoxes(self.p, side.ctypes.data_as(C.POINTER(C.c_int)), _dp(bmin), _dp(bsize))
        _check(n if n < 0 else 0, "smolamdGetBoxes")
        return n, side, bmin, bsize

    def diagnostics(self):
        out = np.zeros(8)
        _check(self.L.smolamdGetDiagnostics(self.p, _dp(out)), "smolamdGetDiagnostics")
        return dict(nmol=int(out[0]), maxedges=int(out[1]), rounds=int(out[2]), err=int(out[3]),
                    serno_next=int(out[4]), step=int(out[5]), molsteps=int(out[6]), rebuilds=int(out[7]))


class Ensemble:
    """Independent simulations stepped together on the device
    (smolamdEnsembleNew): every step is one set of launches for all members,
    and each member gives exactly the results it gives alone.  Members are
    Simulation objects defined but not yet updated or stepped; they must
    stay alive as long as the ensemble."""

    def __init__(self, sims):
        self.L = lib()
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.p for s in self.sims])
        self.p = self.L.smolamdEnsembleNew(arr, len(self.sims))
        if not self.p:
            code, fn, msg = last_error()
            raise SmoldynError("%s: %s (%d)" % (fn, msg, code))

    def __del__(self):
        if getattr(self, "p", None):
            self.L.smolamdEnsembleFree(self.p)
            self.p = None

    def __len__(self):
        return len(self.sims)

    def run_steps(self, n, graph=True):
        ms = C.c_double(0)
        _check(self.L.smolamdEnsembleRunSteps(self.p, n, 1 if graph else 0, C.byref(ms)), "smolamdEnsembleRunSteps")
        return ms.value

    def count(self, member, species=None):
        n = self.L.smolamdEnsembleGetMoleculeCount(self.p, member, species.encode() if species else None)
        _check(n if n < 0 else 0, "smolamdEnsembleGetMoleculeCount")
        return n

    def molecules(self, member, ll=0):
        n = self.L.smolamdEnsembleGetMolecules(self.p, member, ll, None, None, None, None, 0)
        _check(n if n < 0 else 0, "smolamdEnsembleGetMolecules")
        ident = np.zeros(n, np.int32)
        pos = np.zeros((n, 3), np.float64)
        ser = np.zeros(n, np.uint64)
        box = np.zeros(n, np.int32)
        self.L.smolamdEnsembleGetMolecules(self.p, member, ll, ident.ctypes.data_as(C.POINTER(C.c_int)), _dp(pos),
                                           ser.ctypes.data_as(C.POINTER(C.c_ulonglong)),
                                           box.ctypes.data_as(C.POINTER(C.c_int)), n)
        return dict(ident=ident, pos=pos, serno=ser, box=box)

    def diagnostics(self, member):
        out = np.zeros(8)
        _check(self.L.smolamdEnsembleGetDiagnostics(self.p, member, _dp(out)), "smolamdEnsembleGetDiagnostics")
        return dict(nmol=int(out[0]), maxedges=int(out[1]), rounds=int(out[2]), err=int(out[3]),
                    serno_next=int(out[4]), step=int(out[5]), molsteps=int(out[6]))
