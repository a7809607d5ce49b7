"""Particle containers.

``ParticleArray`` has the reference layout (ocelot/cpbd/beam/particle.py:79-84):
``rparticles`` is a (6, n) float64 row-major array x, x', y, y', tau, delta;
``q_array`` the charges [C]; ``E`` the reference energy [GeV]; ``s`` the
position [m].  The drop-in ``SpaceCharge`` accepts this class, the
reference's own ParticleArray, or anything with those attributes.

``DeviceParticleArray`` keeps the same attributes as torch float64 tensors on
a GPU, so a tracking loop can stay in HBM between kicks; ``to_host`` gives a
``ParticleArray`` back.
"""
import numpy as np


class ParticleArray:
    def __init__(self, n=0):
        self.rparticles = np.zeros((6, n))
        self.q_array = np.zeros(n)
        self.s = 0.0
        self.E = 0.0

    @property
    def n(self):
        return self.rparticles.shape[1]

    def size(self):
        return self.n

    def x(self):
        return self.rparticles[0]

    def px(self):
        return self.rparticles[1]

    def y(self):
        return self.rparticles[2]

    def py(self):
        return self.rparticles[3]

    def tau(self):
        return self.rparticles[4]

    def p(self):
        return self.rparticles[5]

    def copy(self):
        c = ParticleArray()
        c.rparticles = self.rparticles.copy()
        c.q_array = self.q_array.copy()
        c.s = self.s
        c.E = self.E
        return c


class DeviceParticleArray:
    """Bunch resident in device memory (torch float64 tensors)."""

    def __init__(self, rparticles, q_array, E, s=0.0):
        import torch
        if not (isinstance(rparticles, torch.Tensor) and rparticles.is_cuda):
            raise TypeError("rparticles must be a CUDA tensor")
        if rparticles.dtype != torch.float64 or rparticles.dim() != 2 or rparticles.shape[0] != 6:
            raise TypeError("rparticles must be a (6, n) float64 tensor")
        if rparticles.stride(1) != 1:
            raise TypeError("rparticles rows must be contiguous")
        if tuple(q_array.shape) != (rparticles.shape[1],):
            raise ValueError(f"q_array must hold one charge per particle: shape {tuple(q_array.shape)}, "
                             f"{rparticles.shape[1]} particles")
        self.rparticles = rparticles
        self.q_array = q_array.to(device=rparticles.device, dtype=torch.float64).contiguous()
        self.E = float(E)
        self.s = float(s)

    @classmethod
    def from_host(cls, p_array, device=0):
        import torch
        r = torch.as_tensor(np.ascontiguousarray(p_array.rparticles, dtype=np.float64)).to(f"cuda:{device}")
        q = torch.as_tensor(np.ascontiguousarray(p_array.q_array, dtype=np.float64)).to(f"cuda:{device}")
        return cls(r, q, p_array.E, getattr(p_array, "s", 0.0))

    @property
    def n(self):
        return self.rparticles.shape[1]

    def to_host(self):
        p = ParticleArray()
        p.rparticles = self.rparticles.cpu().numpy().copy()
        p.q_array = self.q_array.cpu().numpy().copy()
        p.E = self.E
        p.s = self.s
        return p
