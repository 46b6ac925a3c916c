"""Python mirror of the FluidSystem host API (include/fluidsystem.h), same
method names, argument meaning and defaults as the reference class
(libfluid/include/fluidsystem.h:16-126), over the C ABI.

Positions/velocities cross as numpy float32 (N,3) arrays instead of
std::vector<float3>.
"""
from __future__ import annotations

import sys

import numpy as np

from . import _lib as L
from .scene import LfParams, Scene, default_params

DEFAULT_TIMESTEP = 0.01


class FluidSystem:
    def __init__(self):
        self.m_isInit = False
        self._scene: Scene | None = None
        self._params: LfParams = default_params(1)
        self._dirty = True
        self._n = 0

    # -- protected surface of the reference ----------------------------------
    def init(self, num_points: int, res: int) -> None:
        """FluidSystem::init (fluidsystem.cu:196-265)."""
        self.clear()
        p = default_params(res)
        for a in range(3):
            p.gravity[a] = self._params.gravity[a]
        self._params = p
        self._scene = Scene(p, num_points)
        self._n = int(num_points)
        self._dirty = False
        self.m_isInit = True

    def clear(self) -> None:
        if self._scene is not None:
            self._scene.close()
        self._scene = None
        self.m_isInit = False
        self._n = 0

    def createFluidBlock(self, firstPtIdx, numPoints, min_x, min_y, min_z, max_x, max_y, max_z) -> None:
        """FluidSystem::createFluidBlock (fluidsystem.cu:66-127)."""
        if not self.m_isInit:
            return
        if firstPtIdx + numPoints > self._n:
            print(f"FluidSystem::createFluidBlock() - attempting to make {numPoints} points but you only "
                  f"have enough space for {self._n - firstPtIdx}\n!", file=sys.stderr)
            return
        if numPoints == 0:
            return
        if self._dirty:  # the block's mass follows the current rest density (fluidsystem.cu:105)
            self._scene.set_params(self._params)
            self._dirty = False
        # Generated on the device, as the C++ class does (the reference
        # writes the lattice point by point from the host,
        # fluidsystem.cu:113-126); the block sets the mass and particle size
        # (fluidsystem.cu:105,108).
        _, mass, psize = self._scene.setup_fluid_block(numPoints, (min_x, min_y, min_z), (max_x, max_y, max_z),
                                                       first=firstPtIdx)
        self.setMass(mass)
        self.setParticleSize(psize)

    # -- public surface ------------------------------------------------------
    def setup(self, numPoints: int, res: int) -> None:
        self.init(numPoints, res)
        self.setAdhesion(0.0)
        self.setViscosity(0.0)
        self.setSurfaceTension(0.0)
        self.setVDamp(0.5)
        self.createFluidBlock(0, numPoints, 0.1, 0.1, 0.1, 0.9, 0.9, 0.9)

    def setupFromData(self, points, velocity, res: int = 32) -> None:
        """FluidSystem::setupFromData (fluidsystem.cu:179-189): like the
        reference, mass and particle size keep init's defaults of 1."""
        points = np.asarray(points, np.float32).reshape(-1, 3)
        self.init(points.shape[0], res)
        self._scene.upload(points, np.asarray(velocity, np.float32).reshape(-1, 3))

    def exportToData(self):
        """(points, colour, velocity) in the current cell-sorted order."""
        if not self.m_isInit:
            return None
        d = self._scene.download(L.LF_ORDER_CELL)
        return d["pos"], d["colour"], d["vel"]

    def getPscale(self) -> float:
        return float(self._params.particle_size)

    def advance(self, dt: float = DEFAULT_TIMESTEP, substeps: int = 10) -> None:
        if not self.m_isInit:
            return
        if self._dirty:
            self._scene.set_params(self._params)
            self._dirty = False
        self._scene.step(dt, substeps)
        self._scene.sync()

    def _set(self, field, value):
        if getattr(self._params, field) != np.float32(value):
            self._dirty = True
        setattr(self._params, field, value)

    def setRestDensity(self, v): self._set("rest_density", v)
    def setParticleSize(self, v): self._set("particle_size", v)
    def setGasConstant(self, v): self._set("gas_constant", v)
    def setXSPHScale(self, v): self._set("xsph_scale", v)
    def setSurfaceTension(self, v): self._set("surface_tension", v)
    def setViscosity(self, v): self._set("viscosity", v)
    def setAdhesion(self, v): self._set("adhesion", v)
    def setVDamp(self, v): self._set("vdamp", v)
    def setGamma(self, v): self._set("gamma", v)
    def setMass(self, v): self._set("mass", v)

    def setLimits(self, velLimit, accLimit):
        self._set("vel_limit", velLimit)
        self._set("acc_limit", accLimit)

    # -- additions -----------------------------------------------------------
    @property
    def scene(self) -> Scene | None:
        return self._scene

    @property
    def params(self) -> LfParams:
        return self._params
