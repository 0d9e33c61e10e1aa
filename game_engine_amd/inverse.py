"""Inverse rendering through the path-replay adjoint (ge_rt_path_adjoint), as a torch.autograd.Function.

The parameters are the scene's material albedo (M x 3) and roughness (M), its lights' intensities
(L) and the sun's intensity (1). A forward call writes them into the scene's material and light
records and renders the per-pixel mean radiance (H x W x 3, float32) with the path tracer's own
estimator; backward hands dLoss/dI to the adjoint, which replays every path from its counter RNG and
returns dLoss/dparameter summed in float64. There is no tape: the replay recomputes each path.

The estimator's sampling is detached (directions, lobe choices, pdfs and roulette from the
parameters at which backward runs), so the gradient is that of the rendering integral estimated
with those samples: unbiased for the integral's gradient, and equal to central differences of the
same detached estimator.
"""
import ctypes as C

import numpy as np
import torch

from game_engine_amd import abi


class InverseRenderer:
    """A scene and a frame whose material and light parameters a torch optimizer drives."""

    def __init__(self, scene, desc, params):
        if not torch.cuda.is_available():
            raise RuntimeError("inverse rendering runs on the GPU: no HIP device")
        self.scene = scene
        self.params = abi.PathTraceParams.from_buffer_copy(params)
        self.materials = (abi.ResolveMaterial * desc.material_count)()
        C.memmove(self.materials, desc.materials, C.sizeof(self.materials))
        self.lights = (abi.ResolveLight * max(desc.light_count, 1))()
        if desc.light_count:
            C.memmove(self.lights, desc.lights, C.sizeof(abi.ResolveLight) * desc.light_count)
        self.light_count = desc.light_count
        # Float views of the records (albedo at words 0..3 of a material, intensity at word 7 of a
        # light): the parameters go in as whole columns.
        self._material_words = np.frombuffer(self.materials, dtype=np.float32).reshape(desc.material_count, -1)
        self._light_words = np.frombuffer(self.lights, dtype=np.float32).reshape(len(self.lights), -1)
        self.width, self.height = params.width, params.height
        n = self.width * self.height
        self._image = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
        self._d_image = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
        self._coverage = torch.zeros(n, dtype=torch.int32, device="cuda")
        self._gm = torch.zeros((desc.material_count, 4), dtype=torch.float64, device="cuda")
        self._gl = torch.zeros(self.params.light_count + 1, dtype=torch.float64, device="cuda")

    def initial(self):
        """The scene's own parameters: albedo (M, 3), roughness (M), lights (L), sun (1)."""
        albedo = torch.tensor([list(m.albedo[:3]) for m in self.materials], dtype=torch.float32)
        rough = torch.tensor([m.albedo[3] for m in self.materials], dtype=torch.float32)
        lights = torch.tensor([self.lights[i].color_intensity[3] for i in range(self.light_count)], dtype=torch.float32)
        sun = torch.tensor([self.params.sun[3]], dtype=torch.float32)
        return albedo, rough, lights, sun

    def set_parameters(self, albedo, roughness, lights, sun):
        self._material_words[:, 0:3] = albedo.detach().to("cpu", torch.float32).numpy()
        self._material_words[:, 3] = roughness.detach().to("cpu", torch.float32).numpy()
        self.scene.update_materials(0, self.materials)
        if self.light_count:
            self._light_words[:self.light_count, 7] = lights.detach().to("cpu", torch.float32).numpy()
            self.scene.update_lights(0, (abi.ResolveLight * self.light_count).from_buffer(self.lights))
        self.params.sun[3] = float(sun.detach().reshape(-1)[0])

    def render(self):
        """The detached estimator's per-pixel mean (H, W, 3) at the current parameters. At equal
        sampling and scene tables the estimator is the path tracer's, to the bit, so the wavefront
        tracer renders it."""
        self._image.zero_()  # a sample_base past 0 resumes into the accumulator
        torch.cuda.synchronize()  # torch's stream has written what the library's stream reads
        self.scene.path_trace(self.params, self._image.data_ptr(), self._coverage.data_ptr(), device=True)
        self.scene.ctx.synchronize()
        torch.cuda.synchronize()
        img = self._image[:, :3] / float(self.params.spp)
        return img.reshape(self.height, self.width, 3).clone()

    def adjoint(self, grad_image):
        """dLoss/dparameters for dLoss/dI = grad_image (H, W, 3): float64 (M, 4) and (1 + L,)."""
        self._d_image.zero_()
        self._d_image[:, :3] = grad_image.reshape(-1, 3).to(torch.float32)
        torch.cuda.synchronize()  # torch's stream has written what the library's stream reads
        # The forward render at these parameters is the adjoint's primal pass: the library reuses it
        # when nothing has changed since (and traces it again otherwise).
        self.scene.path_adjoint(self.params, self._d_image.data_ptr(), None, None,
                                (self._gm.data_ptr(), self._gl.data_ptr()), device=True, reuse_primal=True)
        torch.cuda.synchronize()
        return self._gm.clone(), self._gl.clone()


class PathReplay(torch.autograd.Function):
    """image = render(albedo, roughness, lights, sun); backward by path replay."""

    @staticmethod
    def forward(ctx, albedo, roughness, lights, sun, renderer):
        renderer.set_parameters(albedo, roughness, lights, sun)
        ctx.renderer = renderer
        ctx.save_for_backward(albedo, roughness, lights, sun)
        return renderer.render().to(albedo.device)

    @staticmethod
    def backward(ctx, grad_image):
        albedo, roughness, lights, sun = ctx.saved_tensors
        r = ctx.renderer
        r.set_parameters(albedo, roughness, lights, sun)  # the parameters the image was rendered at
        gm, gl = r.adjoint(grad_image.to("cuda"))
        dev = albedo.device
        return (gm[:, :3].to(dev, albedo.dtype), gm[:, 3].to(dev, roughness.dtype),
                gl[1:1 + r.light_count].to(dev, lights.dtype), gl[:1].to(dev, sun.dtype), None)


def render(albedo, roughness, lights, sun, renderer):
    return PathReplay.apply(albedo, roughness, lights, sun, renderer)
