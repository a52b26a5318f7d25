"""Host-side mirror of the reference's plugin surface for the path integrators.

The reference registers integrators and renderers by name
(RendererBuilder.AddIntegratorBuilder / AddRendererBuilder,
Config/RendererBuilder.cs:808-846) and picks the renderer through
`integrator.TargetRendererName` (:514-526).  This module provides the same
shapes for the GPU path:

  PathTracer / VolumetricPathTracer   integrator parameters (IIntegrator)
  WavefrontPathIntegrator            TargetRendererName = "wavefront"
  WavefrontRenderer                  Renderer: RenderTarget, Start, Wait,
                                     Stop, IsCompleted, IsSuccess,
                                     AllTaskCount, CompletedTaskCount,
                                     ElapsedTime (Renderer.cs:5-52)
  RendererBuilder                    name -> factory registries

Rendering itself happens in libradiantium_hip.so (rd_render); nothing here
computes radiance.
"""
import ctypes as C
import threading
import time
from datetime import timedelta

import numpy as np

from . import native
from .native import RadiantiumError


class PathTracer:
    """PathTracer parameters (Integrators/PathTracer.cs:26-41), registry
    defaults of "path" (RendererBuilder.cs:56-64)."""
    TargetRendererName = "block_based"

    def __init__(self, max_depth=-1, min_depth=3, rr_threshold=0.99, method="Mis", strategy="Uniform"):
        if method not in ("Bxdf", "Nee", "Mis"):
            raise ValueError(f"unknown method {method}")
        if strategy not in ("Uniform", "All"):
            raise ValueError(f"unknown strategy {strategy}")
        self.MaxDepth, self.MinDepth, self.RRThreshold = max_depth, min_depth, rr_threshold
        self.Method, self.Strategy = method, strategy

    def apply(self, scene_desc):
        scene_desc.set_integrator(type="path", max_depth=self.MaxDepth, min_depth=self.MinDepth,
                                  rr_threshold=self.RRThreshold, method=self.Method, strategy=self.Strategy)


class VolumetricPathTracer(PathTracer):
    """VolumetricPathTracer parameters (Integrators/VolumetricPathTracer.cs)."""

    def __init__(self, max_depth=-1, min_depth=3, rr_threshold=0.99, strategy="Uniform"):
        super().__init__(max_depth, min_depth, rr_threshold, "Mis", strategy)

    def apply(self, scene_desc):
        scene_desc.set_integrator(type="vol_path", max_depth=self.MaxDepth, min_depth=self.MinDepth,
                                  rr_threshold=self.RRThreshold, method="Mis", strategy=self.Strategy)


class WavefrontPathIntegrator:
    """Drop-in integrator: same parameters as PathTracer / VolumetricPathTracer,
    rendered by WavefrontRenderer on the GPU."""
    TargetRendererName = "wavefront"

    def __init__(self, inner):
        self.inner = inner

    def apply(self, scene_desc):
        self.inner.apply(scene_desc)


class WavefrontRenderer:
    """Renderer over rd_render (BlockBasedRenderer's role,
    Renderers/BlockBasedRenderer.cs:8-234).  RenderTarget is the reference's
    column-major ColorBuffer as a float32 array [W*H*3]."""

    def __init__(self, ctx, scene, spp=256, seed=0, wave_paths=0):
        self._ctx = ctx
        self._scene = scene
        self._spp = spp
        self._seed = seed
        self._wave_paths = wave_paths
        self.RenderTarget = np.zeros(scene.width * scene.height * 3, np.float32)
        self.Sum2 = np.zeros_like(self.RenderTarget)
        self._cancel = C.c_int32(0)
        self._progress = C.c_int32(0)
        self._thread = None
        self._error = None
        self._t0 = None
        self._t1 = None
        # one task per pass: a pass holds whole samples, up to 2^27 paths
        npix = scene.width * scene.height
        pass_cap = max(min(1 << 27, npix * spp), npix)
        per_pass = max(1, pass_cap // max(npix, 1))
        self.AllTaskCount = (spp + per_pass - 1) // per_pass

    @property
    def IsCompleted(self):
        return self._thread is not None and not self._thread.is_alive()

    @property
    def IsSuccess(self):
        return self._error is None

    @property
    def Error(self):
        return self._error

    @property
    def CompletedTaskCount(self):
        return self._progress.value

    @property
    def ElapsedTime(self):
        if self._t0 is None:
            return timedelta(0)
        end = self._t1 if self._t1 is not None else time.perf_counter()
        return timedelta(seconds=end - self._t0)

    def _run(self):
        try:
            rgb, s2 = self._scene.render(self._spp, seed=self._seed, wave_paths=self._wave_paths,
                                         cancel=C.pointer(self._cancel), progress=C.pointer(self._progress))
            self.RenderTarget[:] = rgb
            self.Sum2[:] = s2
        except RadiantiumError as e:   # BlockBasedRenderer.cs:161-166: exception -> IsSuccess = false
            self._error = e
        finally:
            self._t1 = time.perf_counter()

    def Start(self):
        if self._thread is not None:
            return self._thread
        self._t0 = time.perf_counter()
        self._thread = threading.Thread(target=self._run, daemon=True)
        self._thread.start()
        return self._thread

    def Wait(self):
        if self._thread is None:
            raise RuntimeError("no render task start")   # BlockBasedRenderer.Wait (:90-94)
        self._thread.join()

    def Stop(self):
        self._cancel.value = 1


class RendererBuilder:
    """Name -> factory registries for the GPU path (RendererBuilder.cs:808-846)."""

    def __init__(self):
        self._integrators = {}
        self._renderers = {}

    def AddIntegratorBuilder(self, name, fn):
        self._integrators[name] = fn

    def AddRendererBuilder(self, name, fn):
        self._renderers[name] = fn

    def CreateIntegrator(self, param):
        if "type" not in param:
            raise ValueError("no key type")
        return self._integrators[param["type"]](self, param)

    def CreateRenderer(self, integrator, ctx, scene, param):
        return self._renderers[integrator.TargetRendererName](self, ctx, scene, param)


def set_default_builders(builder):
    """Registers "wavefront_path" / "wavefront_vol_path" integrators and the
    "wavefront" renderer, with the reference's parameter names and defaults."""

    def path(_, p):
        return WavefrontPathIntegrator(PathTracer(p.get("max_depth", -1), p.get("min_depth", 3),
                                                  p.get("rr_threshold", 0.99), p.get("method", "Mis"),
                                                  p.get("strategy", "Uniform")))

    def vol(_, p):
        return WavefrontPathIntegrator(VolumetricPathTracer(p.get("max_depth", -1), p.get("min_depth", 3),
                                                            p.get("rr_threshold", 0.99),
                                                            p.get("strategy", "Uniform")))

    builder.AddIntegratorBuilder("wavefront_path", path)
    builder.AddIntegratorBuilder("wavefront_vol_path", vol)
    builder.AddRendererBuilder("wavefront", lambda _, ctx, scene, p: WavefrontRenderer(
        ctx, scene, p.get("spp", 256), p.get("seed", 0), p.get("wave_paths", 0)))
    return builder
