"""Drop-in for the reference's `pyphoenix` wrapper (pyphoenix/pyphoenix.py:14-215).

Same class, constructor and methods: `pyphoenix(precision="fp32", use_gpu=True, gpu_id=0,
sfml=False, tetm=False, hint_path=None, result_path=None, doprint=True, debug=False)`,
`check_gpu()`, `get_result_path()`, `find_binary(hint)`, `run(config)`.

Differences, all on the MI355X side:
  - the binary is PHOENIX_gpu_{fp32,fp64} built by phoenix_amd.build (the in-tree
    phoenix_amd/bin is searched first, then hint_path / $PHOENIX_PATH);
  - GPU detection asks the HIP runtime (via libphx) instead of nvidia-smi; the
    device is selected with HIP_VISIBLE_DEVICES;
  - there is no CPU executable and no download fallback: a missing binary is an error.
  - `run_inprocess(config)` runs the same configuration through the C ABI in this
    process and returns the solver handle (no subprocess, fields as numpy arrays).

Additions that the reference wrapper does not have:
  - `pyphoenix.rk4_evolve(P, V, psi0, n0, physics=..., dt=..., steps=..., checkpoint_every=50)`:
    differentiable RK4 evolution (torch.autograd.Function `RK4Adjoint` over the hand-written
    adjoint kernel) giving dJ/dP, dJ/dV, dJ/dpsi0, dJ/dn0 for any objective J of the final state;
  - `pyphoenix.ensemble(sizes, ...)`: uneven grids stepped together, one launch per step;
  - `pyphoenix.physics(N_c, N_r, L_x, L_y)`: the reference's default constants for a grid.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))


class pyphoenix:
    bin_path = None
    have_bin = None
    use_gpu = None
    precision = None
    debug = None
    sfml = None
    tetm = None
    has_gpu = None
    gpu_id = None
    doprint = None

    def __init__(self, precision="fp32", use_gpu=True, gpu_id=0, sfml=False, tetm=False, hint_path=None,
                 result_path=None, doprint=True, debug=False):
        self.precision = precision
        self.use_gpu = use_gpu
        self.debug = debug
        self.sfml = sfml
        self.tetm = tetm
        self.gpu_id = gpu_id
        self.doprint = doprint
        self.has_gpu = self.check_gpu()
        self.have_bin = False
        if use_gpu and not self.has_gpu:
            print("Warning: GPU requested but no HIP device detected; PHOENIX_gpu_* will fail to initialise.")
        if result_path is None:
            os.makedirs("phoenix_tmp", exist_ok=True)
            itmp = 0
            tmpdir = os.path.join("phoenix_tmp", str(itmp))
            while os.path.isdir(tmpdir):
                itmp += 1
                tmpdir = os.path.join("phoenix_tmp", str(itmp))
            self.result_path = os.path.abspath(tmpdir)
        else:
            self.result_path = os.path.abspath(result_path)
        if hint_path is None:
            hint_path = os.environ.get("PHOENIX_PATH", os.path.join(PKG, "bin"))
        self.find_binary(os.path.abspath(hint_path))

    def check_gpu(self) -> bool:
        try:
            import torch  # the HIP runtime PyTorch ships is the one libphx binds to
            return bool(torch.cuda.is_available())
        except Exception:
            return False

    def get_result_path(self):
        return self.result_path

    def find_binary(self, hint_path):
        name = f"PHOENIX_gpu_{self.precision}"
        cands = [os.path.join(PKG, "bin", name)]
        cands += sorted(glob.glob(os.path.join(hint_path, "**", name + "*"), recursive=True))
        for c in cands:
            if os.path.isfile(c) and os.access(c, os.X_OK):
                self.bin_path = os.path.abspath(c)
                self.have_bin = True
                if self.debug:
                    print("Found binary", self.bin_path)
                return True
        self.have_bin = False
        return False

    def _args(self, config):
        """Command line of a config dict as pyphoenix.py:167-186 builds it: one string of
        "--key value ..." split on whitespace, so keys and values may carry spaces (a key
        "initialState " repeats a flag, a value "none gauss+ring" is two tokens)."""
        args = []
        for k in config:
            args += ("--" + str(k)).split()
            values = config[k] if isinstance(config[k], list) else [config[k]]
            for e in values:
                args += str(e).split()
        args += ["--path", self.result_path, "--path", "."]
        if not self.sfml:
            args.append("-nosfml")
        if self.tetm:
            args.append("-tetm")
        return args

    def run(self, config):
        if not self.have_bin:
            raise FileNotFoundError("PHOENIX_gpu_%s is missing: build it with `python -m phoenix_amd.build`" % self.precision)
        os.makedirs(self.result_path, exist_ok=True)
        cmd = [self.bin_path] + self._args(config)
        if self.debug:
            print(" ".join(cmd))
        env = os.environ.copy()
        env["HIP_VISIBLE_DEVICES"] = str(self.gpu_id)
        process = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, env=env, universal_newlines=True)
        if self.doprint:
            # the executable's progress block (printCMD): relayed block by block, as pyphoenix.py:199-210 does
            keys = ("T =", "Progress:", "Current System:", "Runtime:", "Time per ps:")
            block = []
            for line in process.stdout:
                if self.debug:
                    sys.stdout.write(line)
                    continue
                if "Total Runtime" in line:
                    sys.stdout.write(line)
                    continue
                if any(k in line for k in keys):
                    block.append(line)
                if "Time per ps:" in line:
                    sys.stdout.write("".join(block))
                    block = []
        self.stdout, self.stderr = process.communicate()
        if process.returncode != 0:
            print("Something went wrong. Please check the error output:")
            print(self.stderr)
        return process.returncode

    @staticmethod
    def physics(N_c, N_r, L_x, L_y):
        """Reference default physical constants for a grid; returns (phx_physics, magic dt)."""
        from ._lib import physics_defaults
        return physics_defaults(N_c, N_r, L_x, L_y)

    @staticmethod
    def rk4_evolve(pump, pot, psi0, n0, **kw):
        """Differentiable RK4 evolution (see phoenix_amd.adjoint.rk4_evolve)."""
        from .adjoint import rk4_evolve
        return rk4_evolve(pump, pot, psi0, n0, **kw)

    def ensemble(self, sizes, physics, spacing=None, **kw):
        """Batched ensemble of grids of the given (N_c, N_r) sizes (see phoenix_amd._lib.Ensemble)."""
        from ._lib import Ensemble
        return Ensemble(self.precision, physics, sizes, spacing, **kw)

    def run_inprocess(self, config, steps=None):
        """Run `config` through the C ABI in this process; returns the open Solver."""
        from ._lib import Solver
        args = self._args(config)
        s = Solver(args, self.precision)
        if steps is None:
            s.run()
        else:
            s.step(steps)
        return s


def __getattr__(name):
    # torch is imported only when the differentiable path is asked for
    if name in ("RK4Adjoint", "rk4_evolve"):
        from . import adjoint
        return getattr(adjoint, name)
    raise AttributeError(name)
