"""The low-level module the reference's Python package is built on (its
pybind11 extension: source/python/module.cpp:386-1310 over
source/python/Simulation.cpp), restated over libsmoldyn_amd's C-ABI.

Method names, argument order and return values (an ErrorCode for the
setters) follow module.cpp, so reference scripts that call
``smoldyn._smoldyn.Simulation.<method>(s, ...)`` run unchanged.  Display,
colour and lighting calls are accepted and do nothing: the engine has no
graphics.  Every simulation runs on the MI355X engine; nothing here falls
back to a CPU path."""
import ctypes as C
import enum
import os

import numpy as np

import smoldyn_amd as _sa


class ErrorCode(enum.IntEnum):
    """libsmoldyn.h:13-26."""
    ok = 0
    notify = -1
    warning = -2
    nonexist = -3
    all = -4
    missing = -5
    bounds = -6
    syntax = -7
    error = -8
    memory = -9
    bug = -10
    same = -11
    wildcard = -12


class MolecState(enum.IntEnum):
    """smoldyn.h:251-262."""
    soln = 0
    front = 1
    back = 2
    up = 3
    down = 4
    bsoln = 5
    all = 6
    none = 7
    some = 8


class RevParam(enum.IntEnum):
    """smoldyn.h:358-374."""
    none = 0
    irrev = 1
    confspread = 2
    bounce = 3
    pgem = 4
    pgemmax = 5
    pgemmaxw = 6
    ratio = 7
    unbindrad = 8
    pgem2 = 9
    pgemmax2 = 10
    ratio2 = 11
    offset = 12
    fixed = 13


class PanelFace(enum.IntEnum):
    """smoldyn.h:484-490."""
    front = 0
    back = 1
    none = 2
    both = 3


class PanelShape(enum.IntEnum):
    """smoldyn.h:491-501."""
    rect = 0
    tri = 1
    sph = 2
    cyl = 3
    hemi = 4
    disk = 5
    all = 6
    none = 7


class SrfAction(enum.IntEnum):
    """smoldyn.h:503-517."""
    reflect = 0
    trans = 1
    absorb = 2
    jump = 3
    port = 4
    mult = 5
    no = 6
    none = 7
    adsorb = 8
    revdes = 9
    irrevdes = 10
    flip = 11


def _ec(code):
    code = int(code)
    return ErrorCode(code) if code in ErrorCode._value2member_map_ else ErrorCode.error


def _s(x):
    return x.encode() if isinstance(x, str) else x


def _vec(v):
    """A double array for a C pointer argument; None for an empty list (the
    reference passes &vector[0] of an empty vector, which libsmoldyn reads as
    absent)."""
    if v is None or len(v) == 0:
        return None, None
    a = np.ascontiguousarray(v, np.float64)
    return a, _sa._dp(a)


class Simulation:
    """``Simulation(low, high, boundary_type)`` builds a model with smolNewSim
    and sets each dimension's walls from a one- or two-letter type
    (Simulation.cpp:14-32); ``Simulation(path, flags)`` reads and prepares a
    configuration file (Simulation.cpp:34-47)."""

    def __init__(self, *args):
        self._L = _sa.lib()
        self._tmin = 0.0
        self._tmax = 0.0
        self._dt = 0.0
        self._seed = None
        self._accur = 10.0
        self._quitatend = False
        if len(args) >= 1 and isinstance(args[0], (str, os.PathLike)):
            self._sim = _sa.Simulation.from_file(str(args[0]), update=True)
            return
        low, high = list(args[0]), list(args[1])
        btype = list(args[2]) if len(args) > 2 else ["r"] * len(low)
        if len(low) != len(high):
            raise ValueError("dimension mismatch %d != %d" % (len(low), len(high)))
        self._sim = _sa.Simulation(low, high)
        if len(btype) != len(low):
            raise ValueError("one boundary type per dimension is required")
        for d, t in enumerate(btype):
            t = str(t)
            if len(t) == 1:
                self._L.smolSetBoundaryType(self._sim.p, d, -1, _s(t))
            else:
                for side in range(2):
                    self._L.smolSetBoundaryType(self._sim.p, d, side, _s(t[side]))

    # ---- handles and simulation-wide settings (module.cpp:394-520)
    def getSimPtr(self):
        return self._sim.p

    @property
    def simptr(self):
        return self._sim.p

    @property
    def engine(self):
        """The smoldyn_amd.Simulation behind this object."""
        return self._sim

    def setModelpath(self, path):
        return True

    def getBoundaries(self):
        n, side, bmin, bsize = self._sim.boxes()
        return list(bmin), list(bmin + side * bsize)

    def count(self):
        """Model sizes (Simulation.cpp:64-86): dim, species (including the
        empty species), surfaces, compartments and variables."""
        buf = C.create_string_buffer(256)
        nsp = 0
        while True:
            self._L.smolGetSpeciesName(self._sim.p, nsp, buf)
            if not buf.value and nsp > 0:
                self._L.smolClearError()
                break
            nsp += 1
        nsrf = len(self.__dict__.get("_surfaces", ()))
        return {"dim": self._sim.dim, "species": nsp, "surface": nsrf, "compartment": 0, "variables": 0}

    def setSimTimes(self, start, stop, dt):
        self._tmin, self._tmax, self._dt = float(start), float(stop), float(dt)
        return _ec(self._L.smolSetSimTimes(self._sim.p, float(start), float(stop), float(dt)))

    @property
    def start(self):
        return self._tmin

    @start.setter
    def start(self, t):
        self._tmin = float(t)
        self._L.smolSetTimeStart(self._sim.p, float(t))

    @property
    def stop(self):
        return self._tmax

    @stop.setter
    def stop(self, t):
        assert t > 0.0
        self._tmax = float(t)
        self._L.smolSetTimeStop(self._sim.p, float(t))

    @property
    def dt(self):
        return self._dt

    @dt.setter
    def dt(self, dt):
        assert dt > 0.0
        self._dt = float(dt)
        self._L.smolSetTimeStep(self._sim.p, float(dt))

    @property
    def accuracy(self):
        return self._accur

    @accuracy.setter
    def accuracy(self, a):
        # The accuracy scale sets how carefully the reference checks
        # surface-crossing paths; the engine's surface tests are exact.
        self._accur = float(a)

    @property
    def seed(self):
        return self._seed

    @seed.setter
    def seed(self, s):
        self.setRandomSeed(s)

    @property
    def randomSeed(self):
        return self._seed

    @randomSeed.setter
    def randomSeed(self, s):
        self.setRandomSeed(s)

    @property
    def quitatend(self):
        return self._quitatend

    @quitatend.setter
    def quitatend(self, q):
        self._quitatend = bool(q)

    def setFlags(self, flags):
        return ErrorCode.ok

    def setTimeNow(self, t):
        return _ec(self._L.smolSetTimeNow(self._sim.p, float(t)))

    def setTimeStep(self, dt):
        self._dt = float(dt)
        return _ec(self._L.smolSetTimeStep(self._sim.p, float(dt)))

    def setTimeStop(self, t):
        self._tmax = float(t)
        return _ec(self._L.smolSetTimeStop(self._sim.p, float(t)))

    def setRandomSeed(self, seed):
        self._seed = int(seed)
        return _ec(self._L.smolSetRandomSeed(self._sim.p, int(seed)))

    def setPartitions(self, method, value):
        return _ec(self._L.smolSetPartitions(self._sim.p, _s(method), float(value)))

    # ---- display (module.cpp:524-627): accepted, no effect
    def setGraphicsParams(self, *args, **kwargs):
        return ErrorCode.ok

    setTiffParams = setLightParams = setBackgroundStyle = setFrameStyle = setGraphicsParams
    setGridStyle = setTextStyle = addTextDisplay = setMoleculeStyle = setGraphicsParams
    setMoleculeColor = setMoleculeSize = setSurfaceStyle = setGraphicsParams

    def displaySim(self):
        return ErrorCode.ok

    # ---- output files and data (module.cpp:417-442, 628-642)
    def setOutputPath(self, path):
        return _ec(self._L.smolSetOutputPath(self._sim.p, _s(str(path))))

    def addOutputFile(self, filename, suffix, append):
        return _ec(self._L.smolAddOutputFile(self._sim.p, _s(str(filename)), int(suffix), 1 if append else 0))

    def addOutputData(self, dataname):
        return _ec(self._L.smolAddOutputData(self._sim.p, _s(dataname)))

    def getOutputData(self, dataname, erase=True):
        return self._sim.output_data(dataname, bool(erase)).tolist()

    # ---- commands (module.cpp:643-660, Command.cpp:31-78)
    def addCommand(self, cmd, cmd_type, **kwargs):
        """A command of timing type cmd_type; on, off, step and multiplier
        come from kwargs as each type needs them (Command.cpp:42-66)."""
        t = str(cmd_type)[0]
        on = off = step = 0.0
        mult = 1.0
        if t == "@":
            on = float(kwargs["on"])
        elif t in "aAbBeE":
            pass
        elif t in "nN":
            step = float(kwargs["step"])
        elif t in "ixjI":
            on, off, step = float(kwargs["on"]), float(kwargs["off"]), float(kwargs["step"])
            if t == "x":
                mult = float(int(kwargs["multiplier"]))
        else:
            raise ValueError("Command type `%s` is not supported" % t)
        k = self._L.smolAddCommand(self._sim.p, _s(t), on, off, step, mult, _s(cmd))
        if k != 0:
            raise RuntimeError("Failed to add command " + cmd)
        return ErrorCode.ok

    def addCommandStr(self, cmd):
        if self._L.smolAddCommandFromString(self._sim.p, _s(cmd)) != 0:
            raise RuntimeError("Failed to add command " + cmd)
        return ErrorCode.ok

    def runCommand(self, cmd):
        return _ec(self._L.smolRunCommand(self._sim.p, _s(cmd)))

    def readConfigString(self, statement, parameters):
        return _ec(self._L.smolReadConfigString(self._sim.p, _s(statement), _s(parameters)))

    # ---- species and molecules (module.cpp:661-854)
    def addSpecies(self, species, mollist=""):
        return _ec(self._L.smolAddSpecies(self._sim.p, _s(species), _s(mollist) if mollist else None))

    def getSpeciesIndex(self, species):
        return self._L.smolGetSpeciesIndex(self._sim.p, _s(species))

    def speciesIndexNT(self, species):
        return self._L.smolGetSpeciesIndexNT(self._sim.p, _s(species))

    def getSpeciesName(self, index):
        buf = C.create_string_buffer(256)
        self._L.smolGetSpeciesName(self._sim.p, int(index), buf)
        return buf.value.decode()

    def setSpeciesMobility(self, species, state, diffConst, drift=(), difmatrix=()):
        _, dr = _vec(drift)
        _, dm = _vec(difmatrix)
        return _ec(self._L.smolSetSpeciesMobility(self._sim.p, _s(species), int(state), float(diffConst), dr, dm))

    def addMolList(self, mollist):
        return _ec(self._L.smolAddMolList(self._sim.p, _s(mollist)))

    def getMolListIndex(self, mollist):
        return self._L.smolGetMolListIndex(self._sim.p, _s(mollist))

    def molListIndexNT(self, mollist):
        return self._L.smolGetMolListIndexNT(self._sim.p, _s(mollist))

    def setMolList(self, species, state, mollist):
        return _ec(self._L.smolSetMolList(self._sim.p, _s(species), int(state), _s(mollist)))

    def setMaxMolecules(self, n):
        return _ec(self._L.smolSetMaxMolecules(self._sim.p, int(n)))

    def addSolutionMolecules(self, species, number, lowpos=(), highpos=()):
        a, lo = _vec(lowpos)
        b, hi = _vec(highpos)
        return _ec(self._L.smolAddSolutionMolecules(self._sim.p, _s(species), int(number), lo, hi))

    def getMoleculeCount(self, species, state):
        return self._L.smolGetMoleculeCount(self._sim.p, _s(species), int(state))

    # ---- walls and surfaces (module.cpp:855-1075)
    def setBoundaryType(self, dimension, highside, btype):
        return _ec(self._L.smolSetBoundaryType(self._sim.p, int(dimension), int(highside), _s(str(btype))))

    def addSurface(self, surface):
        k = _ec(self._L.smolAddSurface(self._sim.p, _s(surface)))
        if k == ErrorCode.ok:
            self.__dict__.setdefault("_surfaces", []).append(surface)
        return k

    def getSurfaceIndex(self, surface):
        return self._L.smolGetSurfaceIndex(self._sim.p, _s(surface))

    def getSurfaceIndexNT(self, surface):
        return self._L.smolGetSurfaceIndexNT(self._sim.p, _s(surface))

    def setSurfaceAction(self, surface, face, species, state, action, newspecies=""):
        return _ec(self._L.smolSetSurfaceAction(self._sim.p, _s(surface), int(face), _s(species), int(state),
                                                int(action), _s(newspecies) if newspecies else None))

    def setSurfaceRate(self, surface, species, state, state1, state2, rate, newspecies, isinternal):
        return _ec(self._L.smolSetSurfaceRate(self._sim.p, _s(surface), _s(species), int(state), int(state1),
                                              int(state2), float(rate), _s(newspecies) if newspecies else None,
                                              1 if isinternal else 0))

    def addPanel(self, surface, panelshape, panel, axisstring, params):
        prm = np.zeros(max(16, len(params)))
        prm[:len(params)] = params
        return _ec(self._L.smolAddPanel(self._sim.p, _s(surface), int(panelshape), _s(panel) if panel else None,
                                        _s(axisstring) if axisstring else None, _sa._dp(prm)))

    def getPanelIndex(self, surface, panel):
        ps = C.c_int(int(PanelShape.none))
        return self._L.smolGetPanelIndex(self._sim.p, _s(surface), C.byref(ps), _s(panel))

    def getPanelIndexNT(self, surface, panel):
        ps = C.c_int(int(PanelShape.none))
        return self._L.smolGetPanelIndexNT(self._sim.p, _s(surface), C.byref(ps), _s(panel))

    def setPanelJump(self, surface, panel1, face1, panel2, face2, isbidirectional):
        return _ec(self._L.smolSetPanelJump(self._sim.p, _s(surface), _s(panel1), int(face1), _s(panel2), int(face2),
                                            1 if isbidirectional else 0))

    def setSurfaceSimParams(self, parameter, value):
        return _ec(self._L.smolSetSurfaceSimParams(self._sim.p, _s(parameter), float(value)))

    # ---- reactions (module.cpp:1132-1235)
    def addReaction(self, reaction, reactant1, rstate1, reactant2, rstate2, productSpeciesStr, productStates,
                    rate):
        nprd = len(productStates)
        names = (C.c_char_p * max(1, nprd))(*[_s(p) for p in productSpeciesStr])
        states = (C.c_int * max(1, nprd))(*[int(x) for x in productStates])
        return _ec(self._L.smolAddReaction(self._sim.p, _s(reaction), _s(reactant1) if reactant1 else None,
                                           int(rstate1), _s(reactant2) if reactant2 else None, int(rstate2), nprd,
                                           names, states, float(rate)))

    def getReactionIndex(self, order, reaction):
        o = C.c_int(int(order[0]) if len(order) else -1)
        return self._L.smolGetReactionIndex(self._sim.p, C.byref(o), _s(reaction))

    def getReactionIndexNT(self, order, reaction):
        o = C.c_int(int(order[0]) if len(order) else -1)
        return self._L.smolGetReactionIndexNT(self._sim.p, C.byref(o), _s(reaction))

    def setReactionRate(self, reaction, rate, type):
        return _ec(self._L.smolSetReactionRate(self._sim.p, _s(reaction), float(rate), int(type)))

    def setReactionRegion(self, reaction, compartment, surface):
        # Compartment- and surface-restricted reactions are not part of the
        # engine's reaction path.
        return ErrorCode.error

    def setReactionProducts(self, reaction, method, parameter, product, position):
        _, pos = _vec(position)
        return _ec(self._L.smolSetReactionProducts(self._sim.p, _s(reaction), int(method), float(parameter),
                                                   _s(product) if product else None, pos))

    # ---- running (Simulation.cpp:157-227)
    def updateSim(self):
        return _ec(self._L.smolUpdateSim(self._sim.p))

    def runSim(self, stoptime, dt, display=True, overwrite=False):
        """Open the output files, set the time step and stop time, update
        and run (Simulation.cpp:157-194)."""
        L, p = self._L, self._sim.p
        er = L.smolOpenOutputFiles(p, 1 if overwrite else 0)
        if er != 0:
            return _ec(er)
        er1 = L.smolSetTimeStep(p, float(dt))
        er2 = L.smolSetTimeStop(p, float(stoptime))
        if er1 != 0 or er2 != 0:
            return _ec(er)
        self._dt, self._tmax = float(dt), float(stoptime)
        er = L.smolUpdateSim(p)
        if er != 0:
            return _ec(er)
        return _ec(L.smolRunSim(p))

    def runUntil(self, stop, dt, display=True, overwrite=False):
        """Simulation.cpp:196-221: output files opened, the time step set when
        dt > 0, the model updated, then run until the break time."""
        L, p = self._L, self._sim.p
        L.smolOpenOutputFiles(p, 1 if overwrite else 0)
        if dt > 0.0:
            self._dt = float(dt)
            L.smolSetTimeStep(p, float(dt))
        L.smolUpdateSim(p)
        return _ec(L.smolRunSimUntil(p, float(stop)))

    def runTimeStep(self):
        return _ec(self._L.smolRunTimeStep(self._sim.p))

    def connect(self, func, target, step, args):
        raise NotImplementedError("Python callbacks during a run are not supported by the device engine")


def getVersion():
    return _sa.lib().smolGetVersion()


def getError(clear=False):
    code, fn, msg = _sa.last_error()
    if clear:
        _sa.lib().smolClearError()
    return _ec(code), msg


def setDebugMode(mode):
    _sa.lib().smolSetDebugMode(1 if mode else 0)


def errorCodeToString(code):
    buf = C.create_string_buffer(64)
    return _sa.lib().smolErrorCodeToString(int(code), buf).decode()
