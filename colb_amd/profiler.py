"""Profiller (utils/data_collector.py:4-55): per-(frame, Newton iteration) residual and line-search
step arrays, saved as the same npz (keys ``convergence`` and ``alphas``).

The reference's `step` overwrites rows iter..15 after each frame through an off-by-one frame
index (abd.py:701-702); this recorder writes row `frame` only and leaves unused iterations 0.
"""
import numpy as np


class Profiller:
    def __init__(self, max_iters, frames=4000):
        self.convergence = np.zeros((frames, max_iters))
        self.alphas = np.zeros_like(self.convergence)

    def recorded_attributes(self):
        return {"convergence": float, "alphas": float}

    def record(self, frame, residuals, alphas):
        if 0 <= frame < self.convergence.shape[0]:
            n = min(len(residuals), self.convergence.shape[1])
            self.convergence[frame, :n] = residuals[:n]
            self.alphas[frame, :n] = alphas[:n]

    def save(self, filename):
        np.savez(filename, **{k: getattr(self, k) for k in self.recorded_attributes()})

    def load(self, filename):
        data = np.load(filename)
        for k in self.recorded_attributes():
            if k in data:
                setattr(self, k, data[k])
