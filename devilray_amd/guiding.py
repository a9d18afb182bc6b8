"""Progressive rendering with online guide training (host driver over the C ABI).

A pass renders `spp_per_pass` samples per pixel into one device-resident FP64
accumulation buffer; while training, the same pass collects training samples
(eligible Lambert vertices, thinned to fit the guide's stage; those that
carried radiance enter the guide's training ring) and `train_steps` Adam
steps on the ring follow on the device. Every pass's samples are unbiased
estimates whatever the state of the network, so all passes are accumulated
into the image. The plain sampler is the same loop with no guide.

The collection fraction is sized so that the stage (4 x the guide's
capacity) does not fill: the first collecting pass uses a fraction at which
even paths that all reach max_depth would offer 3/4 of the stage, and the
passes after it a fraction that offers 3/4 of the stage at the first pass's
measured rate. The ring is filled in sample order, so with a stage that
never fills, training on the same scene and seeds gives the same network
bit for bit.
"""
from __future__ import annotations

import numpy as np

from . import DR_GUIDE_COLLECT, DR_GUIDE_USE, DrCamera, DrRenderParams, Guide, Scene


class ProgressiveRenderer:
    def __init__(self, scene: Scene, camera: DrCamera, guide: Guide | None = None, spp_per_pass: int = 4,
                 max_depth: int = 12, rr_depth: int = 3, seed: int = 0, train_passes: int | None = None,
                 train_steps: int = 1, lr_schedule=None, guide_after: int = 1, stream=None):
        self.scene, self.camera, self.guide = scene, camera, guide
        self.spp_per_pass = spp_per_pass
        self.max_depth, self.rr_depth, self.seed = max_depth, rr_depth, seed
        self.train_passes = train_passes  # None: train on every pass
        self.train_steps = train_steps
        self.lr_schedule = lr_schedule  # None, or pass index -> Adam step size for that pass's steps
        self.guide_after = max(1, guide_after)  # training passes before the guide samples (warm-up)
        self.stream = stream
        self.npix = camera.width * camera.height
        self.accum = scene.ctx.alloc(self.npix * 32)
        self.accum.copy_from(np.zeros((self.npix, 4), np.float64))
        self.passes = 0
        self.trained = 0
        self._fraction_set = False
        self._fraction = 1.0

    def _training(self) -> bool:
        return self.guide is not None and (self.train_passes is None or self.passes < self.train_passes)

    def step(self):
        """Render one pass (and train once while training). Asynchronous,
        except for one synchronisation after the first collecting pass that
        sizes the collection fraction."""
        p = DrRenderParams(self.spp_per_pass, self.passes * self.spp_per_pass, self.max_depth, self.rr_depth,
                           self.seed, 0, 0)
        if self.guide is None:
            self.scene.render_dev(self.camera, p, self.accum.ptr, self.stream)
        else:
            flags = DR_GUIDE_USE if self.trained >= self.guide_after else 0
            train = self._training()
            if train:
                flags |= DR_GUIDE_COLLECT
            stage = 4 * self.guide.capacity
            if train and not self._fraction_set:
                # at most max_depth - 1 vertices of a path are eligible
                bound = self.npix * self.spp_per_pass * max(1, self.max_depth - 1)
                self._fraction = min(1.0, 0.75 * stage / bound)
                self.guide.set_options(self._fraction, self.guide.guided_vertices, self.guide.train_batch)
            self.scene.render_guided_dev(self.camera, p, self.guide, flags, self.accum.ptr, self.stream)
            if train:
                if not self._fraction_set:
                    _, total = self.guide.copy_samples(0)
                    if total > 0:
                        self._fraction = min(1.0, 0.75 * stage * self._fraction / total)
                        self.guide.set_options(self._fraction, self.guide.guided_vertices, self.guide.train_batch)
                        self._fraction_set = True
                if self.lr_schedule is not None:
                    self.guide.set_learning_rate(float(self.lr_schedule(self.passes)))
                for _ in range(self.train_steps):
                    self.guide.train_step_dev(stream=self.stream)
                self.trained += 1
        self.passes += 1

    def image(self) -> np.ndarray:
        """Mean radiance per pixel (height, width, 3), FP64."""
        self.scene.ctx.synchronize()
        a = self.accum.to_numpy(np.float64, self.npix * 4).reshape(self.camera.height, self.camera.width, 4)
        return a[..., :3] / np.maximum(a[..., 3:4], 1.0)

    def close(self):
        self.accum.free()
