"""Headless frame loop of the reference application (src/renderer.cpp,
include/renderer.hpp, include/timer.hpp): the same scene set-up and the
clear -> input -> update -> draw -> present shape, with the Win32 window
replaced by recorded camera input and a present callback (for instance
`png_presenter`, which writes each frame as PNG).

    r = Renderer()
    r.init(1240, 720, model="assets/Gas Tank/Gas Tank.obj")
    times = r.run(frames=120, present=png_presenter("out/frame_%04d.png"))

Differences from the reference, all of them harness-side: the model's
rotation angle is 0.4 rad/s times a frame clock (frame index x dt) instead
of the wall clock (GetTickCount64, renderer.cpp:105), so a run is
reproducible; input comes from an `InputState` sequence instead of the
window; timing is returned (and averaged like DeltaTimeAverager) instead of
printed.
"""
from __future__ import annotations

import time
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import render as R

BLACK = R.Color(0, 0, 0, 255)


def _mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """a * b for column-major float32 4x4 (16,) arrays, each column of the
    result summed ((a0 b.x + a1 b.y) + a2 b.z) + a3 b.w in float32."""
    A = a.reshape(4, 4)
    B = b.reshape(4, 4)
    out = np.empty((4, 4), np.float32)
    for c in range(4):
        acc = A[0] * B[c, 0]
        acc = acc + A[1] * B[c, 1]
        acc = acc + A[2] * B[c, 2]
        out[c] = acc + A[3] * B[c, 3]
    return out.reshape(16)


def translate(v) -> np.ndarray:
    """glm::translate(mat4(1), v), column-major."""
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = np.asarray(v, np.float32)
    return m.reshape(16)


def rotate(angle: float, axis) -> np.ndarray:
    """glm::rotate(mat4(1), angle, axis) in float32 (axis normalized first)."""
    a = np.float32(angle)
    c, s = np.float32(np.cos(a)), np.float32(np.sin(a))
    ax = np.asarray(axis, np.float32)
    ax = ax * np.float32(1.0 / np.sqrt(np.float32(np.dot(ax, ax))))
    t = (np.float32(1) - c) * ax
    m = np.eye(4, dtype=np.float32)
    m[0, :3] = [c + t[0] * ax[0], t[0] * ax[1] + s * ax[2], t[0] * ax[2] - s * ax[1]]
    m[1, :3] = [t[1] * ax[0] - s * ax[2], c + t[1] * ax[1], t[1] * ax[2] + s * ax[0]]
    m[2, :3] = [t[2] * ax[0] + s * ax[1], t[2] * ax[1] - s * ax[0], c + t[2] * ax[2]]
    return m.reshape(16)


def model_matrix(time_s: float) -> np.ndarray:
    """reference Renderer::render (renderer.cpp:100-110): translate(0, -3, 3)
    times a rotation of 0.4 * t about +y."""
    return _mul(translate((0.0, -3.0, 3.0)), rotate(0.4 * time_s, (0.0, 1.0, 0.0)))


class DeltaTimeAverager:
    """reference DeltaTimeAverager (include/timer.hpp:73-113): the mean frame
    time over each `interval` seconds, collected instead of printed."""

    def __init__(self, interval: float = 1.0, description: str = "Average Delta Time"):
        self.interval, self.description = interval, description
        self.reports: List[float] = []      # ms per frame, one per interval
        self._acc, self._n, self._since = 0.0, 0, 0.0

    def addDelta(self, dt: float) -> None:
        self._acc += dt
        self._n += 1
        self._since += dt
        if self._since >= self.interval:
            self.reports.append(self._acc * 1000.0 / self._n if self._n else 0.0)
            self._acc, self._n, self._since = 0.0, 0, 0.0


def png_presenter(pattern: str) -> Callable[[int, R.Framebuffer], None]:
    """Present by writing frame i to pattern % i (PNG, top row first)."""
    def present(i: int, fb: R.Framebuffer) -> None:
        fb.write_png(pattern % i)
    return present


class Renderer:
    """reference Renderer (include/renderer.hpp, src/renderer.cpp:58-133)."""

    def __init__(self, ctx: Optional[R.Context] = None):
        self.ctx = ctx
        self._own_ctx = ctx is None
        self.m_Meshes: List[R.Mesh] = []
        self.m_Shaders: List[R.IShader] = []

    def init(self, screenWidth: int, screenHeight: int, model: Optional[str] = None,
             filter: int = R.FILTER_NEAREST) -> None:
        """renderer.cpp:58-94: framebuffer, camera at (0, 0, 5) looking at the
        origin (fov 60), a tiled pipeline, the three shaders with PBR current,
        and the model (an OBJ path; None: no mesh yet, see `addMesh`)."""
        if self.ctx is None:
            self.ctx = R.Context(0)
        self.m_Framebuffer = R.Framebuffer(self.ctx, screenWidth, screenHeight)
        self.m_Camera = R.Camera((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), 60.0, screenWidth / screenHeight)
        self.m_Camera.setViewport(0, 0, screenWidth, screenHeight)
        self.m_DefaultPipeline = R.TiledPipeline(self.m_Camera, self.m_Framebuffer)
        self.m_Shaders = [R.FlatShader(lightDirection=(0.0, -1.0, 0.0)),
                          R.PhongShader(lightDirection=(0.0, -1.0, 0.0), lightColor=(0.6, 0.6, 0.6)),
                          R.PBRShader(lightDirection=(0.0, -1.0, 0.0), lightColor=(5.6, 5.6, 5.6))]
        self.m_CurrentShader = self.m_Shaders[2]
        self.m_DefaultPipeline.setShader(self.m_CurrentShader)
        if model is not None:
            self.m_Meshes.append(R.Mesh.from_obj(self.ctx, model, filter=filter))

    def addMesh(self, mesh: R.Mesh) -> None:
        self.m_Meshes.append(mesh)

    def setShader(self, index: int) -> None:
        self.m_CurrentShader = self.m_Shaders[index]
        self.m_DefaultPipeline.setShader(self.m_CurrentShader)

    def drawMesh(self, model, mesh: R.Mesh) -> None:
        self.m_DefaultPipeline.drawMesh(model, mesh, asynchronous=True)

    def render(self, time_s: float) -> None:
        """renderer.cpp:100-110: every mesh at translate(0,-3,3) * rotate(0.4 t)."""
        model = model_matrix(time_s)
        for mesh in self.m_Meshes:
            self.drawMesh(model, mesh)

    def run(self, frames: int, dt: float = 1.0 / 60.0, inputs: Optional[Sequence] = None,
            present: Optional[Callable[[int, R.Framebuffer], None]] = None) -> List[float]:
        """renderer.cpp:112-133 for `frames` frames: input, camera update,
        clear, render, present.  inputs[i] is frame i's window state (an
        InputState or anything with its three calls); time is i * dt.
        Returns the wall time of every frame in seconds (a frame is waited
        for before its time is taken)."""
        times = []
        self.averager = DeltaTimeAverager(1.0, "Averaged Delta Time")
        for i in range(frames):
            t0 = time.perf_counter()
            if inputs is not None:
                self.m_Camera.handleInput(inputs[i])
            self.m_Camera.update(dt)
            self.m_Framebuffer.clearColor(BLACK)
            self.m_Framebuffer.clearDepth()
            self.render(i * dt)
            if present is not None:
                present(i, self.m_Framebuffer)
            self.ctx.sync()
            times.append(time.perf_counter() - t0)
            self.averager.addDelta(times[-1])
        return times

    def close(self) -> None:
        for m in self.m_Meshes:
            m.close()
        self.m_Meshes = []
        if getattr(self, "m_Framebuffer", None) is not None:
            self.m_Framebuffer.close()
            self.m_Framebuffer = None
        if self._own_ctx and self.ctx is not None:
            self.ctx.close()
            self.ctx = None
