"""The run configurations of the benchmark and tests (box 4x1x1, Nmvp = 0, Fx = Fy = 0).

The reference keeps c with the displacing fluid at c = 0 and
mu = exp(Vratio * c) (pllcart3d.c:502, 853); a front written as
c' = 0.5 erfc((x-1)/w) with mu = exp(R (1-c')) is the same state with
c = 1 - c'.  The solver works in the reference's convention.
"""
from ._lib import IC_ERF, IC_MULTIMODE, IC_NOISE, Params

# name: (nx, ny, nz, Vratio R, Pe, front width, IC recipe, seed, amplitude, steps)
CONFIGS = {
    # plain FP64 reference case, 64x32x32, R=2, Pe=500, w=0.05, +-0.01 noise (seed 1), 200 steps
    "noise_64x32x32": dict(nx=64, ny=32, nz=32, Vratio=2.0, Pe=500.0, Inf_thick=0.05,
                 ic=IC_NOISE, seed=1, amp=0.01, steps=200),
    # 256x128x128 (4.2M cells), R=3, Pe=2000, w=0.02, 16x16-mode front (seed 2), 1000 steps
    "front_256x128x128": dict(nx=256, ny=128, nz=128, Vratio=3.0, Pe=2000.0, Inf_thick=0.02,
                 ic=IC_MULTIMODE, seed=2, amp=0.01, steps=1000),
    # 512x256x256 (33.5M cells), Pe=4000, same front recipe, 500 steps
    "front_512x256x256": dict(nx=512, ny=256, nz=256, Vratio=3.0, Pe=4000.0, Inf_thick=0.02,
                 ic=IC_MULTIMODE, seed=2, amp=0.01, steps=500),
}


def make_params(nx, ny, nz, Vratio=2.0, Pe=500.0, Inf_thick=0.05, Re=1.0, Inf_pos=1.0,
                CFLmax=0.3, Timestep=2e-3, X=(0.0, 4.0), Y=(-0.5, 0.5), Z=(-0.5, 0.5), Fx=0.0, Fy=0.0,
                Nmvp=0, Mum=1.5, Lambda=0.0):
    return Params(nx, ny, nz, X[0], X[1], Y[0], Y[1], Z[0], Z[1], Re, Pe, Vratio, Fx, Fy,
                  Inf_pos, Inf_thick, CFLmax, Timestep, Nmvp, Mum, Lambda)


def config_params(name):
    c = CONFIGS[name]
    return make_params(c["nx"], c["ny"], c["nz"], Vratio=c["Vratio"], Pe=c["Pe"], Inf_thick=c["Inf_thick"])


# the batched ensemble: members drawn with numpy's PCG64 from seed 1000 + m:
# grid uniform from three shapes, R ~ U[1, 4], Pe log-uniform in [200, 4000],
# a +-0.01 noise front (recipe NOISE, seed 1000 + m); 400 steps each
ENSEMBLE_SHAPES = [(64, 32, 32), (128, 48, 48), (192, 64, 64)]


def ensemble_members(n=64, seed0=1000):
    """[(params, ic_recipe, seed, amp)] of the first n members."""
    import numpy as np
    out = []
    for m in range(n):
        rng = np.random.default_rng(seed0 + m)
        nx, ny, nz = ENSEMBLE_SHAPES[int(rng.integers(0, len(ENSEMBLE_SHAPES)))]
        R = float(rng.uniform(1.0, 4.0))
        Pe = float(np.exp(rng.uniform(np.log(200.0), np.log(4000.0))))
        out.append((make_params(nx, ny, nz, Vratio=R, Pe=Pe, Inf_thick=0.05), IC_NOISE, seed0 + m, 0.01))
    return out
