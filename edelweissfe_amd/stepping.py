"""Steps, increments and step actions of the nonlinear implicit static path.

Restates, with the reference's names and semantics:
  * TimeStep                     edelweissfe/timesteppers/timestep.py
  * IncrementController          the increment logic of edelweissfe/timesteppers/adaptivetimestepper.py:60-183
                                 (zero-size first increment, growth x1.1 after 3 good increments, cutback,
                                 limits), as an iterator of its own design
  * AdaptiveStep                 edelweissfe/steps/adaptivestep.py:56-156 (defaults and options)
  * Dirichlet step action        edelweissfe/stepactions/dirichlet.py:53-158 (delta per node set,
                                 linear amplitude, components)
  * NodeForces step action       edelweissfe/stepactions/nodeforces.py:53-152
  * exceptions                   edelweissfe/utils/exceptions.py:36-73
"""

from __future__ import annotations

import numpy as np


class CutbackRequest(Exception):
    def __init__(self, message, cutbackSize):
        super().__init__(message)
        self.cutbackSize = cutbackSize


class ReachedMaxIterations(Exception):
    pass


class DivergingSolution(Exception):
    pass


class ReachedMaxIncrements(Exception):
    pass


class ReachedMinIncrementSize(Exception):
    pass


class StepFailed(Exception):
    pass


class ConditionalStop(Exception):
    pass


class TimeStep:
    """One increment as the solver sees it (edelweissfe/timesteppers/timestep.py): its number in the step, its
    size and the step progress at its end (both fractions of the step), and the matching times."""

    __slots__ = ("number", "stepProgressIncrement", "stepProgress", "timeIncrement", "stepTime", "totalTime")

    def __init__(self, number, stepProgressIncrement, stepProgress, timeIncrement, stepTime, totalTime):
        self.number = number
        self.stepProgressIncrement = stepProgressIncrement
        self.stepProgress = stepProgress
        self.timeIncrement = timeIncrement
        self.stepTime = stepTime
        self.totalTime = totalTime


class IncrementController:
    """Adaptive increment sizes of one step, with the behaviour of the reference's AdaptiveTimeStepper
    (edelweissfe/timesteppers/adaptivetimestepper.py:90-183):

    * the first increment handed out has size 0 (it only evaluates the start state);
    * sizes are fractions of the step, clipped to what is left of it;
    * once three increments in a row have been handed out since the last discard, and the solver did not veto
      growth for the last one (``holdSize``), every request grows the size by ``growth`` up to ``largest``;
    * ``discard(f)`` takes back the last increment and shrinks the size to f times itself, clamped to
      [smallest, largest]; discarding the zero-size increment, or shrinking a size already at ``smallest``,
      raises ReachedMinIncrementSize;
    * asking for another increment after more than ``maxCount`` raises ReachedMaxIncrements.

    Iterating over the controller yields TimeStep objects; the bookkeeping of an increment (counting it,
    growing the size) happens when the next one is requested, so a discard in between acts on the increment
    just handed out."""

    def __init__(self, startTime, stepLength, firstSize, largest, smallest, maxCount, growth=1.1):
        self._t0 = startTime
        self._length = stepLength
        self._largest = largest
        self._smallest = smallest
        self._maxCount = maxCount
        self._growth = growth
        self._size = min(firstSize, largest)  # size of the next nonzero increment
        self._done = 0.0                      # step progress reached by the increments handed out
        self._count = 0                       # number of the increment handed out last
        self._streak = 0                      # increments handed out since the start or the last discard
        self._vetoGrowth = False
        self._started = False

    def __iter__(self):
        return self

    def _settleLast(self):
        """Bookkeeping of the increment handed out last, done when the next one is requested."""
        if self._count > self._maxCount:
            raise ReachedMaxIncrements()
        if self._streak >= 3 and not self._vetoGrowth:
            self._size = min(self._size * self._growth, self._largest)
        self._vetoGrowth = False
        self._count += 1
        self._streak += 1

    def __next__(self):
        if self._started:
            self._settleLast()
        if not self._done < 1.0 - 1e-15:
            raise StopIteration
        self._started = True
        self._size = min(self._size, 1.0 - self._done)
        size = self._size if self._count > 0 else 0.0
        self._done += size
        inStep = self._length * self._done
        return TimeStep(self._count, size, self._done, self._length * size, inStep, self._t0 + inStep)

    def holdSize(self):
        """The next request must not grow the size (slow convergence of the increment just solved)."""
        self._vetoGrowth = True

    def discard(self, factor):
        """Takes back the increment handed out last and shrinks the size by `factor`."""
        if self._count == 0:
            raise ReachedMinIncrementSize()
        self._done -= self._size
        self._count -= 1
        self._streak = 0
        if self._size == self._smallest:
            raise ReachedMinIncrementSize()
        self._size = min(max(self._size * factor, self._smallest), self._largest)


class AdaptiveStep:
    """A step with its actions; ``actions`` maps action type -> {name: action} like the reference."""

    defaultStartInc = 1.0
    defaultMaxInc = 1.0
    defaultMinInc = 1e-4
    defaultMaxNumInc = 1000
    defaultMaxIter = 10
    defaultCriticalIter = 5
    defaultMaxGrowingIter = 10
    defaultCutbackFactor = 0.25

    def __init__(self, number, model, stepActions=None, **kwargs):
        kw = {k.lower(): v for k, v in kwargs.items()}
        self.number = number
        self.model = model
        self.length = float(kw.get("steplength", 1.0))
        self.startIncrementSize = float(kw.get("startinc", self.defaultStartInc))
        self.maxIncrementSize = float(kw.get("maxinc", self.defaultMaxInc))
        self.minIncrementSize = float(kw.get("mininc", self.defaultMinInc))
        self.maxNumberIncrements = int(kw.get("maxnuminc", self.defaultMaxNumInc))
        self.maxIter = int(kw.get("maxiter", self.defaultMaxIter))
        self.criticalIter = int(kw.get("criticaliter", self.defaultCriticalIter))
        self.maxGrowIter = int(kw.get("maxgrowiter", self.defaultMaxGrowingIter))
        self.cutbackFactor = float(kw.get("cutbackfactor", self.defaultCutbackFactor))
        self.incrementGenerator = IncrementController(model.time, self.length, self.startIncrementSize,
                                                      self.maxIncrementSize, self.minIncrementSize,
                                                      self.maxNumberIncrements)
        self.actions = {k: {} for k in ("dirichlet", "nodeforces", "distributedload", "bodyforce", "geostatic",
                                        "options", "modelupdate")}
        for typ, acts in (stepActions or {}).items():
            self.actions.setdefault(typ, {}).update(acts)

    def getTimeStep(self):
        return iter(self.incrementGenerator)

    def discardAndChangeIncrement(self, cutbackFactor):
        return self.incrementGenerator.discard(cutbackFactor)

    def preventIncrementIncrease(self):
        return self.incrementGenerator.holdSize()


def _linear(x):
    return x


class Dirichlet:
    """``dirichlet, nSet=.., field=displacement, 1=.., 2=.., 3=..`` (dirichlet.py:53-158). ``components``
    maps 0-based component -> prescribed increment over the step; ``amplitude`` is f(stepProgress); ``scale``
    (one value per node of the set, the ``analyticalField`` option) multiplies the values node by node. A later
    step redefines it with ``updateStepAction``: the new values are the next step's increments."""

    def __init__(self, name, model, nSet: str, components: dict, field="displacement", amplitude=None, scale=None):
        if field != "displacement":
            raise NotImplementedError("the GPU path carries the displacement field only")
        self.name = name
        self.field = field
        self.nSetName = nSet
        self.nSet = np.asarray(model.nodeSets[nSet], dtype=np.int64)
        self.updateStepAction(components, amplitude, scale)

    def updateStepAction(self, components: dict, amplitude=None, scale=None):
        comps = sorted(int(c) for c in components)
        self._components = comps
        self.delta = np.tile([float(components[c]) for c in comps], (len(self.nSet), 1))
        if scale is not None:
            self.delta = self.delta * np.asarray(scale, dtype=np.float64).reshape(-1, 1)
        self.amplitude = amplitude or _linear
        self.active = True

    @property
    def components(self):
        return self._components

    def applyAtStepStart(self, model):
        pass

    def applyAtIncrementStart(self, model, timeStep):
        pass

    def applyAtStepEnd(self, model):
        self.active = False

    def getDelta(self, timeStep: TimeStep):
        if self.active:
            return self.delta * (self.amplitude(timeStep.stepProgress) -
                                 self.amplitude(timeStep.stepProgress - timeStep.stepProgressIncrement))
        return self.delta * 0.0


class NodeForces:
    """``nodeforces, nSet=.., components='fx, fy, fz'`` (nodeforces.py:53-152): a force vector per node,
    scaled by the amplitude of the step progress."""

    def __init__(self, name, model, nSet: str, vector, field="displacement", amplitude=None):
        if field != "displacement":
            raise NotImplementedError("the GPU path carries the displacement field only")
        self.name = name
        self.field = field
        self.nSetName = nSet
        self.nodeSet = np.asarray(model.nodeSets[nSet], dtype=np.int64)
        self.nodeForcesStepStart = np.zeros((len(self.nodeSet), 3))
        self.nodeForcesDelta = np.tile(np.asarray(vector, dtype=np.float64), (len(self.nodeSet), 1))
        self.amplitude = amplitude or _linear
        self._idle = False

    def applyAtStepStart(self, model):
        pass

    def applyAtIncrementStart(self, model, timeStep):
        pass

    def applyAtStepEnd(self, model):
        if not self._idle:
            self.nodeForcesStepStart += self.nodeForcesDelta * self.amplitude(1.0)
            self.nodeForcesDelta[:] = 0
            self._idle = True

    def updateStepAction(self, vector, amplitude=None):
        """nodeforces.py:85-108: the new vector is added over the next step, on top of the loads so far."""
        self.nodeForcesDelta = np.tile(np.asarray(vector, dtype=np.float64), (len(self.nodeSet), 1))
        self.amplitude = amplitude or _linear
        self._idle = False

    def getCurrentLoad(self, timeStep: TimeStep):
        if self._idle:
            return self.nodeForcesStepStart
        return self.nodeForcesStepStart + self.nodeForcesDelta * self.amplitude(timeStep.stepProgress)


class BodyForce:
    """``bodyforce, elSet=.., forceVector='bx, by, bz'`` (edelweissfe/stepactions/bodyforce.py:49-116): a force
    per unit volume on an element set, held at its step-end value afterwards; ``amplitude`` is f(stepProgress).
    A later step changes it with ``updateStepAction(forceVector=...)`` (total) or ``(delta=...)``."""

    def __init__(self, name, model, elSet: str, vector, amplitude=None):
        self.name = name
        if elSet not in model.elementSets:
            raise KeyError(f"BodyForce {name}: unknown element set {elSet}")
        self.elSetName = elSet
        self.blockIndex, self.elements = model.elementSets[elSet]
        load = np.asarray(vector, dtype=np.float64).ravel()
        if len(load) < model.domainSize:
            raise ValueError(f"BodyForce {name}: force vector has wrong dimension!")
        self.forceAtStepStart = np.zeros(model.domainSize)
        self.delta = load[: model.domainSize]
        self.amplitude = amplitude or _linear
        self._idle = False

    def applyAtStepStart(self, model):
        pass

    def applyAtIncrementStart(self, model, timeStep):
        pass

    def applyAtStepEnd(self, model):
        if not self._idle:
            self.forceAtStepStart = self.forceAtStepStart + self.delta * self.amplitude(1.0)
            self.delta = self.delta * 0.0
            self._idle = True

    def updateStepAction(self, forceVector=None, delta=None, amplitude=None):
        if forceVector is not None:
            self.delta = np.asarray(forceVector, dtype=np.float64).ravel()[:3] - self.forceAtStepStart
        elif delta is not None:
            self.delta = np.asarray(delta, dtype=np.float64).ravel()[:3]
        self.amplitude = amplitude or _linear
        self._idle = False

    def getCurrentLoad(self, timeStep: TimeStep):
        t = 1.0 if self._idle else timeStep.stepProgress
        return self.forceAtStepStart + self.delta * self.amplitude(t)

    @property
    def elementSet(self):
        return self.elements


class DistributedLoad:
    """``distributedload, surface=.., type=pressure, magnitude=p`` (edelweissfe/stepactions/distributedload.py:
    52-120): a load on the faces of an element-based surface, held at its step-end value afterwards; ``amplitude``
    is f(stepProgress). A later step changes it with ``updateStepAction(magnitude=...)`` (total) or
    ``(delta=...)``; the GPU path carries the load type pressure."""

    def __init__(self, name, model, surface: str, loadType: str, magnitude, amplitude=None):
        self.name = name
        if surface not in model.surfaces:
            raise KeyError(f"DistributedLoad {name}: unknown surface {surface}")
        self.surfaceName = surface
        self._surface = {}
        for face, elSet in model.surfaces[surface].items():
            self._surface[int(face)] = model.elementSets[elSet] if isinstance(elSet, str) else elSet
        self._loadType = loadType
        self.magnitudeAtStepStart = 0.0
        self.delta = np.asarray(magnitude, dtype=np.float64).ravel()
        self.amplitude = amplitude or _linear
        self._idle = False

    @property
    def surface(self):
        """{face number: (block index, element indices)}."""
        return self._surface

    @property
    def loadType(self):
        return self._loadType

    def applyAtStepStart(self, model):
        pass

    def applyAtIncrementStart(self, model, timeStep):
        pass

    def applyAtStepEnd(self, model):
        if not self._idle:
            self.magnitudeAtStepStart = self.magnitudeAtStepStart + self.delta * self.amplitude(1.0)
            self.delta = 0
            self._idle = True

    def updateStepAction(self, magnitude=None, delta=None, amplitude=None):
        if magnitude is not None:
            self.delta = np.asarray(magnitude, dtype=np.float64).ravel() - self.magnitudeAtStepStart
        elif delta is not None:
            self.delta = np.asarray(delta, dtype=np.float64).ravel()
        self.amplitude = amplitude or _linear
        self._idle = False

    def getCurrentLoad(self, timeStep: TimeStep):
        t = 1.0 if self._idle else timeStep.stepProgress
        return self.magnitudeAtStepStart + self.delta * self.amplitude(t)


class SetField:
    """``setfield, fieldOutput=.., type=uniform|analyticalField, value=..`` (edelweissfe/stepactions/setfield.py):
    at the start of the step, a per-element quadrature-point result of the field output (its element set and
    points) is set to a value, or to an analytical field at the points. On the GPU path the results are those of
    the small-strain J2 material: ``kappa`` (state index 6) and ``stress``. The reference integrates the J2 stress
    incrementally from its stored value; the GPU path stores the plastic strain instead, so a prescribed stress
    s is stored as the plastic strain eps - C^-1 s at the current strain eps, which gives the same stress and the
    same trial stresses afterwards."""

    RESULTS = {("vonmises", "kappa"): 1, ("vonmises", "stress"): 6}  # values per point

    def __init__(self, name, model, result: str, elSet: str, quadraturePoints, kind: str, value):
        self.name = name
        if kind not in ("uniform", "analyticalField"):
            raise Exception(f"Invalid type: {kind}")
        self.result = result
        self.blockIndex, self.elements = model.elementSets[elSet]
        self.quadraturePoints = quadraturePoints  # slice or index array
        self.kind = kind
        self.value = value  # list of floats (uniform) or f(x, y, z)
        self.active = True

    def applyAtStepEnd(self, model):
        self.active = False

    def updateStepAction(self):
        self.active = True

    def applyAtIncrementStart(self, model, timeStep):
        pass

    def qpValues(self, model):
        """(element indices, quadrature point indices, values [nElements, nPoints, nComponents]) to be set."""
        from .model import quadraturePointCoordinates

        block = model.elementBlocks[self.blockIndex]
        qps = np.arange(block.nInt)[self.quadraturePoints]
        if self.kind == "uniform":
            v = np.asarray(self.value, dtype=np.float64)
            return self.elements, qps, np.broadcast_to(v, (len(self.elements), len(qps), len(v))).copy()
        xq = quadraturePointCoordinates(block, model.nodeCoordinates)[self.elements][:, qps]
        vals = np.array([[self.value(*x) for x in row] for row in xq], dtype=np.float64)
        return self.elements, qps, vals.reshape(len(self.elements), len(qps), 1)

    def applyAtStepStart(self, model):
        if not self.active:
            return
        from .model import ELEMENT_LIBRARY, quadraturePointStrains

        block = model.elementBlocks[self.blockIndex]
        key = (block.material.name if block.material else None, self.result)
        if (key not in self.RESULTS or block.stateVars is None or block.planOrder is None
                or ELEMENT_LIBRARY[block.elType]["tl"]):
            raise NotImplementedError(f"setfield of {self.result} for {key[0]} is not part of the GPU path")
        els, qps, vals = self.qpValues(model)
        if vals.shape[2] != self.RESULTS[key]:
            raise ValueError(f"SetField {self.name}: {self.result} has {self.RESULTS[key]} components per point")
        els = np.asarray(els)
        nEl, nInt = len(block), block.nInt
        st = block.stateVars.detach().cpu().numpy().reshape(-1, nInt, nEl)  # [state][point][plan position]
        pos = np.empty(nEl, dtype=np.int64)
        pos[np.asarray(block.planOrder)] = np.arange(nEl)
        if self.result == "kappa":
            for j, q in enumerate(qps):
                st[6, q, pos[els]] = vals[:, j, 0]
        else:  # stress s -> plastic strain eps - C^-1 s (engineering shears)
            used = np.unique(np.concatenate([b.conn.ravel() for b in model.elementBlocks]))
            Un = np.zeros((model.nNodes, 3))
            U = getattr(model, "U", None)
            if U is not None:
                Un[used] = U.detach().cpu().numpy().reshape(-1, 3)
            eps = quadraturePointStrains(block, model.nodeCoordinates, Un)[els][:, qps]  # [e, q, 6]
            mp = block.elementMaterialProperties()[els]
            E, nu = mp[:, 0, None], mp[:, 1, None]
            s = vals
            e_el = np.empty_like(s)
            for i in range(3):
                e_el[..., i] = (s[..., i] - nu * (s[..., (i + 1) % 3] + s[..., (i + 2) % 3])) / E
            e_el[..., 3:] = s[..., 3:] * (2 * (1 + nu) / E)[:, :, None]
            ep = eps - e_el
            for j, q in enumerate(qps):
                st[0:6, q, pos[els]] = ep[:, j, :].T
        block.stateVars.copy_(block.stateVars.new_tensor(st.ravel()))


class Options:
    """``options, category=NISTSolver, key=value`` (edelweissfe/stepactions/options.py:38-61)."""

    def __init__(self, name, category, **options):
        self.name = name
        self.category = category
        self.options = options

    def applyAtStepStart(self, model):
        pass

    def applyAtIncrementStart(self, model, timeStep):
        pass

    def applyAtStepEnd(self, model):
        pass
