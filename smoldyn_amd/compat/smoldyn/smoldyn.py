"""The reference's user-level Python API (source/python/smoldyn/smoldyn.py)
over the MI355X engine.

Class names, constructor arguments and the calls they make on the
low-level simulation follow the reference: ``Species`` (smoldyn.py:84-294),
the panels and ``Surface`` (smoldyn.py:312-1181), ``Reaction`` and
``BidirectionalReaction`` (smoldyn.py:1277-1606), the partitions
(smoldyn.py:1608-1643) and ``Simulation`` (smoldyn.py:1645-2410).  A model
script written for the reference runs on the device engine when this
package's directory comes first on ``sys.path``.  Features outside the
engine (compartments, ports, lattices, Python callbacks, 1-D and 2-D
models) raise an error at the call or at the first update; display settings
are accepted and have no effect."""
import functools
import operator
import warnings
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Union

from . import _smoldyn

__all__ = ["__version__", "version", "getError", "Simulation", "Species", "Panel", "Triangle", "Rectangle",
           "Sphere", "Hemisphere", "Cylinder", "Disk", "Surface", "Partition", "MoleculePerBox", "Box",
           "Reaction", "BidirectionalReaction", "NullSpecies"]

__version__ = "%g" % _smoldyn.getVersion()


def version():
    return __version__


def getError(clear=False):
    return _smoldyn.getError(clear)


def _toMS(st):
    """A state name or MolecState as a MolecState (smoldyn.py:72-81)."""
    return _smoldyn.MolecState.__members__[st] if isinstance(st, str) else _smoldyn.MolecState(int(st))


def _ok(k, what):
    if k != _smoldyn.ErrorCode.ok:
        code, msg = _smoldyn.getError()
        raise RuntimeError("%s failed (%s): %s" % (what, k, msg))


class Species:
    """A species in one state (smoldyn.py:84-294): addSpecies, then the
    mobility of each state in ``difc`` and the molecule list."""

    def __init__(self, simulation, name, state="soln", color=None, difc=0.0, display_size=3, mol_list=""):
        self.simulation = simulation
        self.name = name
        assert self.name
        self._difc = {}
        self._mol_list = ""
        self._color = color if color is not None else {"soln": "black"}
        self._displaySize = display_size
        _ok(self.simulation.addSpecies(self.name), "addSpecies")
        if state not in _smoldyn.MolecState.__members__:
            raise NameError("%s is not a valid MolecState" % state)
        self.state = state
        self.difc = difc if isinstance(difc, dict) else {self.state: difc}
        if mol_list:
            self.mol_list = mol_list

    def __repr__(self):
        return "<Species: %s state=%s difc=%s>" % (self.name, self.state, self._difc)

    def setStyle(self, **kwargs):
        """Display style; no effect on the device engine."""

    @property
    def difc(self):
        return self._difc

    @difc.setter
    def difc(self, difconst):
        self._difc = {self.state: difconst} if not isinstance(difconst, dict) else difconst
        for state, D in self._difc.items():
            _ok(self.simulation.setSpeciesMobility(self.name, _toMS(state), D), "setSpeciesMobility")

    @property
    def mol_list(self):
        return self._mol_list

    @mol_list.setter
    def mol_list(self, val):
        _ok(self.simulation.addMolList(val), "addMolList")
        _ok(self.simulation.setMolList(self.name, _toMS(self.state), val), "setMolList")
        self._mol_list = val

    @property
    def color(self):
        return self._color

    @color.setter
    def color(self, c):
        self._color = c

    @property
    def display_size(self):
        return self._displaySize

    @display_size.setter
    def display_size(self, s):
        self._displaySize = s

    def addToSolution(self, N, pos=(), highpos=(), lowpos=()):
        """N molecules at pos, or uniformly between lowpos and highpos, or
        uniformly in the system (smoldyn.py:263-294)."""
        assert self.state == "soln", "You can't use this function on a Species with type %s" % self.state
        if pos:
            lowpos = highpos = pos
        _ok(self.simulation.addSolutionMolecules(self.name, N, list(lowpos), list(highpos)), "addSolutionMolecules")


class NullSpecies(Species):
    """The absent reactant of a zero- or first-order reaction."""

    def __init__(self):
        self.name = ""
        self.state = "none"

    def __bool__(self):
        return False

    def __len__(self):
        return 0


# ---- panels and surfaces (smoldyn.py:312-1181)
class Panel:
    def __init__(self, *, name="", shape=_smoldyn.PanelShape.none, neighbors=(), simulation=None):
        self.axisstr = ""
        self.name = name
        self.ctype = _smoldyn.PanelShape(shape)
        self.pts = []
        self.simulation = simulation
        self.front = _PanelFace(name="front", panel=self, simulation=simulation)
        self.back = _PanelFace(name="back", panel=self, simulation=simulation)
        self.surface = None
        if neighbors:
            raise NotImplementedError("panel neighbors are not supported by the device engine")

    def _setSimulation(self, simulation):
        self.simulation = simulation
        self.front.simulation = simulation
        self.back.simulation = simulation

    def _getName(self, index):
        return self.name if self.name else "%s%d" % (self.ctype.name, index)

    def jumpTo(self, face1, panel2, face2, bidirectional=False):
        return getattr(self, face1).jumpTo(getattr(panel2, face2), bidirectional)

    def __str__(self):
        return "<Panel %s %s %s>" % (self.ctype.name, self.name, self.pts)


class Rectangle(Panel):
    def __init__(self, corner, dimensions, axis, name="", simulation=None):
        super().__init__(simulation=simulation, shape=_smoldyn.PanelShape.rect, name=name)
        assert axis[0] in "+-", "axis must precede by '+' or '-'"
        assert axis[1].lower() in "012xyz"
        self.corner, self.dimensions = corner, dimensions
        self.axisstr = axis.lower()
        self.pts = [*corner, *dimensions]


class Triangle(Panel):
    def __init__(self, *, vertices=((),), name="", simulation=None):
        super().__init__(simulation=simulation, shape=_smoldyn.PanelShape.tri, name=name)
        self.vertices = vertices
        self.pts = functools.reduce(operator.iconcat, [list(v) for v in vertices], [])


class Sphere(Panel):
    def __init__(self, *, center, radius, slices, stacks=0, name="", simulation=None):
        super().__init__(simulation=simulation, shape=_smoldyn.PanelShape.sph, name=name)
        assert slices > 0, "Positive int is required"
        self.center, self.radius, self.slices = center, radius, slices
        self.stacks = stacks if stacks > 0 else slices
        self.pts = [*center, radius, slices, self.stacks]


class Hemisphere(Panel):
    def __init__(self, *, center, radius, vector, slices, stacks, name="", simulation=None):
        super().__init__(simulation=simulation, shape=_smoldyn.PanelShape.hemi, name=name)
        self.center, self.radius, self.vector, self.slices, self.stacks = center, radius, vector, slices, stacks
        self.pts = [*center, radius, *vector, slices, stacks]


class Cylinder(Panel):
    def __init__(self, *, start, end, radius, slices, stacks, name="", simulation=None):
        super().__init__(simulation=simulation, shape=_smoldyn.PanelShape.cyl, name=name)
        self.start, self.end, self.radius, self.slices, self.stacks = start, end, radius, slices, stacks
        self.pts = [*start, *end, radius, slices, stacks]


class Disk(Panel):
    def __init__(self, *, center, radius, slices, vector, name="", simulation=None):
        super().__init__(simulation=simulation, shape=_smoldyn.PanelShape.disk, name=name)
        assert slices > -1, "Slices must be positive. Given value %s" % slices
        self.center, self.radius, self.slices, self.vector = center, radius, slices, vector
        self.pts = [*center, radius, *vector, slices]


class _PanelFace:
    def __init__(self, *, simulation, name, panel):
        assert name in ("front", "back")
        self.simulation = simulation
        self.name = _smoldyn.PanelFace.__members__[name]
        self.panel = panel

    def jumpTo(self, toface, bidirectional=False):
        """smoldyn.py:697-724: a jump between faces of two panels of one
        surface and one shape."""
        assert self.panel.ctype == toface.panel.ctype, "Both panels must have same geometry"
        assert self.panel.surface is toface.panel.surface, "Both panels must have same surface"
        _ok(self.simulation.setPanelJump(self.panel.surface.name, self.panel.name, self.name, toface.panel.name,
                                         toface.name, bidirectional), "setPanelJump")


class _SurfaceFaceCollection:
    _actions = ("reflect", "trans", "absorb", "jump", "port", "mult", "no", "none")

    def __init__(self, simulation, faces, surfname):
        self.simulation = simulation
        self.faces = faces
        self.surfname = surfname

    def setStyle(self, *args, **kwargs):
        """Drawing style; no effect on the device engine."""

    def setAction(self, species, action, new_species=None):
        """smoldyn.py:802-860: the action for each species (objects, not
        names) on each face of the collection."""
        assert action in self._actions, "'%s' is not a valid action" % action
        if isinstance(species, str):
            raise NameError("This API does not accepts species by their names or the keyword 'all'."
                            " Please pass the objects.")
        if new_species:
            raise NotImplementedError("This feature is not implemented.")
        for sp in ([species] if isinstance(species, Species) else species):
            sname, sstate = (sp.name, sp.state) if isinstance(sp, Species) else (sp, "soln")
            for which in self.faces:
                _ok(self.simulation.setSurfaceAction(self.surfname, _smoldyn.PanelFace.__members__[which], sname,
                                                     _toMS(sstate), _smoldyn.SrfAction.__members__[action]),
                    "setSurfaceAction")


class Surface:
    """A named surface with its panels (smoldyn.py:922-1181)."""

    def __init__(self, *, simulation, name, panels):
        self.simulation = simulation
        self.panels = panels
        self.name = name
        if self.name:
            _ok(self.simulation.addSurface(self.name), "addSurface")
            self.front = _SurfaceFaceCollection(simulation, ["front"], name)
            self.back = _SurfaceFaceCollection(simulation, ["back"], name)
            self.both = _SurfaceFaceCollection(simulation, ["front", "back"], name)
            for i, panel in enumerate(self.panels):
                panel.name = panel._getName(i)
                panel.surface = self
                panel._setSimulation(simulation)
                _ok(simulation.addPanel(name, panel.ctype, panel.name, panel.axisstr, panel.pts), "addPanel")

    def setStyle(self, face, *args, **kwargs):
        """Drawing style; no effect on the device engine."""

    def setAction(self, face, species, action, new_species=None):
        getattr(self, face).setAction(species, action, new_species)

    def _setRate(self, species, state1, state2, rate, new_species=None, isinternal=False):
        if isinstance(species, Species):
            sname, sstate = species.name, species.state
        else:
            sname, sstate = species[0].name, species[1]
        _ok(self.simulation.setSurfaceRate(self.name, sname, _toMS(sstate), _toMS(state1), _toMS(state2), rate,
                                           new_species.name if new_species else "", isinternal), "setSurfaceRate")

    def setRate(self, species, state1, state2, rate, revrate=0.0, new_species=None, isinternal=False):
        """smoldyn.py:1141-1181: the rate from state1 to state2, and the
        reverse rate when revrate > 0."""
        self._setRate(species, state1, state2, rate, new_species, isinternal)
        if revrate > 0.0:
            self._setRate(species, state2, state1, revrate, new_species, isinternal)

    def addMolecules(self, *args, **kwargs):
        raise NotImplementedError("surface-bound molecules are not supported by the device engine")


# ---- reactions (smoldyn.py:1277-1606)
class Reaction:
    __methoddict__ = dict(i="irrev", p="pgem", x="pgemmax", y="pgemmax2", r="ratio", b="unbindrad", q="pgem2",
                          s="ratio2", o="offset", f="fixed")

    def __init__(self, simulation, name, subs, prds, *, rate=0.0, compartment=None, surface=None,
                 binding_radius=0.0, reaction_probability=0.0):
        self.simulation = simulation
        self.name = name
        self.__rate = 0.0
        self.subs = subs
        self.prds = prds
        self.reaction_probability = reaction_probability
        self.binding_radius = binding_radius
        self.compartment = compartment
        self.surface = surface
        assert len(subs) < 3, "At most two reactants are supported."
        if subs is None or len(subs) == 0:
            assert len(prds) > 0, "At least one product for a zero-order reaction."
            subs = [NullSpecies()]
        r1 = subs[0]
        r2 = subs[1] if len(subs) == 2 else NullSpecies()
        r1name, r1state = (r1.name, _toMS(r1.state)) if isinstance(r1, Species) else (r1[0].name, _toMS(r1[1]))
        r2name, r2state = (r2.name, _toMS(r2.state)) if isinstance(r2, Species) else (r2[0].name, _toMS(r2[1]))
        prdNames, prdStates = [], []
        for x in prds:
            if isinstance(x, Species):
                prdNames.append(x.name)
                prdStates.append(_toMS(x.state))
            else:
                prdNames.append(x[0].name)
                prdStates.append(_toMS(x[1]))
        k = self.simulation.addReaction(name, r1name, r1state, r2name, r2state, prdNames, prdStates, rate)
        if k != _smoldyn.ErrorCode.ok:
            raise RuntimeError("Failed to add reaction, ErrorCode %s" % k)
        self.setRate(rate, reaction_probability, binding_radius)
        if compartment is not None or surface is not None:
            raise NotImplementedError("compartment- and surface-restricted reactions are not supported "
                                      "by the device engine")

    @property
    def rate(self):
        return self.__rate

    @rate.setter
    def rate(self, rate):
        if rate != self.__rate:
            self.__rate = rate
            self.setRate(rate)

    def setRate(self, rate, reaction_probability=None, binding_radius=-1.0):
        """smoldyn.py:1400-1430: a rate >= 0 sets the rate constant;
        otherwise the reaction probability (first order), or the probability
        and binding radius (second order)."""
        if rate >= 0.0:
            _ok(self.simulation.setReactionRate(self.name, rate, False), "setReactionRate")
            return
        if len(self.subs) < 2:
            assert reaction_probability is not None, "Must set rate or reaction_probability"
            _ok(self.simulation.setReactionRate(self.name, reaction_probability, True), "setReactionRate")
        else:
            if reaction_probability is not None:
                assert 0.0 <= reaction_probability <= 1.0, "reaction_probability must be between 0 and 1"
                _ok(self.simulation.setReactionRate(self.name, reaction_probability, 2), "setReactionRate")
            if binding_radius >= 0.0:
                _ok(self.simulation.setReactionRate(self.name, binding_radius, True), "setReactionRate")

    @property
    def order(self):
        return len(self.subs)

    def setIntersurface(self, rules):
        raise NotImplementedError("intersurface reactions are not supported by the device engine")

    def productPlacement(self, method, param=0.0, product="", pos=()):
        """smoldyn.py:1469-1515: placement method by name or one-letter code."""
        method = self.__methoddict__.get(method, method)
        revType = _smoldyn.RevParam.__members__[method]
        if method in ("fixed", "offset"):
            assert product, "Product is required"
            assert pos, "pos is required"
        _ok(self.simulation.setReactionProducts(self.name, revType, param, product, list(pos)),
            "setReactionProducts")


class BidirectionalReaction:
    """A forward reaction and, when kb > 0, its reverse, named name+'fwd'
    and name+'rev' (smoldyn.py:1517-1606)."""

    def __init__(self, simulation, name, subs, prds, kf, kb=0.0, *, binding_radius=0.0, reaction_probability=0.0,
                 compartment=None, surface=None):
        self.simulation = simulation
        self.name = name
        self._kf, self._kb = kf, kb
        fwdname, revname = (name + "fwd", name + "rev") if kb > 0.0 else (name, "")
        self.forward = Reaction(simulation, fwdname, subs, prds, rate=kf, compartment=compartment, surface=surface,
                                binding_radius=binding_radius, reaction_probability=reaction_probability)
        self.reverse = None
        if kb > 0.0:
            self.reverse = Reaction(simulation, revname, prds, subs, rate=kb, compartment=compartment,
                                    surface=surface)

    @property
    def kf(self):
        return self._kf

    @kf.setter
    def kf(self, val):
        self._kf = val
        self.forward.setRate(val)

    @property
    def kb(self):
        return self._kb

    @kb.setter
    def kb(self, val):
        self._kb = val
        self.reverse.setRate(val)


# ---- partitions (smoldyn.py:1608-1643)
class Partition:
    def __init__(self, simulation, name, value):
        self.simulation, self.name, self.value = simulation, name, value
        _ok(self.simulation.setPartitions(self.name, self.value), "setPartitions")


class MoleculePerBox(Partition):
    def __init__(self, simulation, size):
        warnings.warn("MoleculePerBox is deprecated. Please use setPartitions in the future", DeprecationWarning)
        super().__init__(simulation, "molperbox", size)


class Box(Partition):
    def __init__(self, simulation, size):
        warnings.warn("Box is deprecated. Please use setPartitions in the future", DeprecationWarning)
        super().__init__(simulation, "boxsize", size)


class Simulation(_smoldyn.Simulation):
    """A complete model (smoldyn.py:1645-2410)."""

    def __init__(self, low, high, *, boundary_type="r", types=None, log_level=3, accuracy=None, output_files=(),
                 seed=-1, quit_at_end=False):
        assert low, "You must pass a valid low bound, current value %s" % (low,)
        assert high, "You must pass a valid high bound, current value %s" % (high,)
        if types is not None:
            warnings.warn("types is deprecated. Please use boundary_type", DeprecationWarning)
            boundary_type = types
        if isinstance(boundary_type, str):
            if len(boundary_type) == 1:
                assert len(low) == len(high), "dimension mismatch %d != %d" % (len(low), len(high))
                boundary_type = boundary_type * len(low)
            boundary_type = list(boundary_type)
        super().__init__(low, high, boundary_type)
        if accuracy:
            self.accuracy = accuracy
        self.setOutputFiles(list(output_files))
        if seed >= 0:
            self.randomSeed = seed
        self.quitatend = quit_at_end

    @classmethod
    def fromFile(cls, path, arg=""):
        """A simulation read and prepared from a configuration file."""
        obj = cls.__new__(cls)
        _smoldyn.Simulation.__init__(obj, str(Path(path).resolve()), arg)
        return obj

    def setOutputFiles(self, outfiles, append=True):
        assert isinstance(outfiles, (list, tuple)), "Expecting a list of files"
        for outfile in outfiles:
            self.setOutputFile(outfile, append)

    def setOutputFile(self, outfile, append=False):
        """smoldyn.py:1783-1803: the file's directory becomes the output path."""
        outfile = Path(outfile).resolve()
        if outfile.parent != Path("."):
            _ok(super().setOutputPath(str(outfile.parent) + "/"), "setOutputPath")
        _ok(super().addOutputFile(outfile.name, 0, append), "addOutputFile")

    def setGraphics(self, *args, **kwargs):
        """Graphics; no effect on the device engine."""

    def setText(self, *args, **kwargs):
        """Text display; no effect on the device engine."""

    def setTiff(self, *args, **kwargs):
        """TIFF output; no effect on the device engine."""

    def setLight(self, *args, **kwargs):
        """Lighting; no effect on the device engine."""

    def run(self, stop, dt, *, start=0.0, display=True, overwrite=False, accuracy=None, log_level=3,
            quit_at_end=False):
        """setSimTimes(start, stop, dt), then open the output files, update
        and run to stop (smoldyn.py:1947-1996)."""
        self.stop = float(stop)
        self.dt = float(dt)
        self.start = float(start)
        if accuracy is not None:
            self.accuracy = float(accuracy)
        super().setSimTimes(self.start, self.stop, self.dt)
        self.quitatend = quit_at_end
        assert self.dt > 0.0, "dt can't be <= 0.0! dt=%s" % self.dt
        assert self.stop > 0.0, "stop time can't be <= 0.0! stop=%s" % self.stop
        k = super().runSim(self.stop, self.dt, display, overwrite)
        if k != _smoldyn.ErrorCode.ok:
            code, msg = _smoldyn.getError()
            raise AssertionError("Expected ErrorCode.ok, got %s: %s" % (k, msg))

    def runUntil(self, stop, dt, display=True, overwrite=False):
        self.stop = stop
        self.dt = float(dt)
        assert self.dt > 0.0, "dt can't be <= 0.0! dt=%s" % self.dt
        assert self.stop > 0.0, "stop time can't be <= 0.0! stop=%s" % self.stop
        return super().runUntil(self.stop, self.dt, display, overwrite)

    def updateSim(self):
        return super().updateSim()

    def display(self):
        return super().displaySim()

    def addCommand(self, cmd, cmd_type, **kwargs):
        """smoldyn.py:2031-2061: step defaults to the time step."""
        if "step" not in kwargs:
            kwargs["step"] = self.dt
        return super().addCommand(cmd, cmd_type[0], **kwargs)

    def addCommandStr(self, cmd):
        return super().addCommandStr(cmd)

    def addSpecies(self, name, state="soln", color=None, difc=0.0, display_size=3, mol_list=""):
        return Species(super(), name, state, color, difc, display_size, mol_list)

    def addMolecules(self, species, number, pos=(), lowpos=(), highpos=()):
        species.addToSolution(number, pos, lowpos=lowpos, highpos=highpos)

    def addSurface(self, name, *, panels):
        return Surface(simulation=super(), name=name, panels=panels)

    def addBidirectionalReaction(self, name, subs, prds, kf, kb=0.0, *, binding_radius=0.0, reaction_probability=0.0,
                                 compartment=None, surface=None):
        return BidirectionalReaction(super(), name, subs, prds, kf, kb, binding_radius=binding_radius,
                                     reaction_probability=reaction_probability, compartment=compartment,
                                     surface=surface)

    def addReaction(self, name, subs, prds, *, rate=0.0, compartment=None, surface=None, binding_radius=0.0,
                    reaction_probability=0.0):
        return Reaction(super(), name, subs, prds, rate=rate, compartment=compartment, surface=surface,
                        binding_radius=binding_radius, reaction_probability=reaction_probability)

    def addPartition(self, name, value):
        return Partition(super(), name, value)

    def addBox(self, size):
        return Box(super(), size)

    def addMoleculePerBox(self, size):
        return MoleculePerBox(super(), size)

    def addCompartment(self, name, *, surface, point):
        raise NotImplementedError("compartments are not supported by the device engine")

    def addPort(self, *args, **kwargs):
        raise NotImplementedError("ports are not supported by the device engine")

    def addOutputData(self, dataname):
        assert " " not in dataname, "spaces not allowed in dataname: '%s'" % dataname
        _ok(super().addOutputData(dataname), "addOutputData")

    def getOutputData(self, dataname, erase=True):
        assert " " not in dataname, "Must not contain spaces: '%s'" % dataname
        return super().getOutputData(dataname, erase)

    def addPanel(self, panel):
        return panel

    def addTriangle(self, *args, **kwargs):
        return Triangle(*args, **kwargs)

    def addRectangle(self, *args, **kwargs):
        return Rectangle(*args, **kwargs)

    def addSphere(self, *args, **kwargs):
        return Sphere(*args, **kwargs)

    def addHemisphere(self, *args, **kwargs):
        return Hemisphere(*args, **kwargs)

    def addCylinder(self, *args, **kwargs):
        return Cylinder(*args, **kwargs)

    def addDisk(self, *args, **kwargs):
        return Disk(*args, **kwargs)
