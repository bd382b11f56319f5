"""Material-parameter identification from reaction-force histories.

The gradient comes from the library's hand-written adjoint
(wfx_adjoint_forward / wfx_adjoint_backward: checkpoint every K steps,
recompute each segment, reverse step kernels).  ``ReactionHistory`` exposes
the forward solve as a torch.autograd.Function of the parameter vector
theta = (sigma_y0, Et, hg_coeff) for the bilinear law, (K, m, hg_coeff)
for the Hollomon law (its plateau stress sigma_y0 held fixed), (A, B,
hg_coeff) for the Johnson-Cook law (n, C, eps_0, m and the uniform
temperature held fixed) or (C1, n2, hg_coeff) for the GMT law (its other
constants, ranges and the uniform temperature held fixed), so a loss written
in torch differentiates through the whole explicit solve.
"""
import numpy as np

from . import workloads
from .domain import WFX_BILINEAR, WFX_FP64, WFX_HOLLOMON, Domain

PARAMS = ("sy0", "Et", "hg_coeff")
LAW_PARAMS = {"bilinear": PARAMS, "hollomon": ("K", "m", "hg_coeff"), "johnson_cook": ("A", "B", "hg_coeff"),
              "gmt": ("C1", "n2", "hg_coeff")}
# Johnson-Cook constants held fixed while A and B vary: hardening exponent n,
# rate constant C at reference rate eps_0, softening exponent m between the
# transition and melting temperatures, and the specimen's uniform temperature
JC_FIXED = dict(n=0.26, C=0.014, eps_0=1.0, m=1.03, T_m=1793.0, T_t=293.0, temp=500.0)
# GMT constants held fixed while C1 and n2 vary: n1, C2, m1, m2, I1, I2, the
# ranges of plastic strain, strain rate and temperature, and the specimen's
# uniform temperature
GMT_FIXED = dict(n1=-2e-4, C2=-1.5e-3, m1=0.0, m2=0.02, I1=-1e-5, I2=1e-4, e_min=0.001, e_max=3.0, er_min=1e-3,
                 er_max=1e5, T_min=20.0, T_max=1200.0, temp=300.0)


class CompressionProblem:
    """Compression with the bottom face fixed in z and the top face moving
    down: by default a (cells)^3 reduced-integration hexa cube of edge
    `size` (or any mesh given as (x, elnod, nodxelem, bottom, top)), nsteps
    steps at dt = 0.8 x the critical step.  The reaction force of the top
    face is the observed history.  law "bilinear": theta = (sigma_y0, Et,
    hg_coeff); "hollomon": theta = (K, m, hg_coeff) with sigma_y0 from mat;
    "johnson_cook": theta = (A, B, hg_coeff) with the constants jc (default
    JC_FIXED); "gmt": theta = (C1, n2, hg_coeff) with the constants gmt
    (default GMT_FIXED).  cfl > 0: the per-step CFL time step with that
    factor, starting from the fixed step."""

    def __init__(self, cells=30, size=1.0, v_top=-1.0, nsteps=2000, every=50, precision=WFX_FP64,
                 mat=workloads.STEEL, stream=0, mesh=None, law="bilinear", jc=None, gmt=None, cfl=0.0):
        if law not in LAW_PARAMS:
            raise ValueError(f"law must be one of {sorted(LAW_PARAMS)}")
        self.law, self.params = law, LAW_PARAMS[law]
        self.jc = dict(JC_FIXED if jc is None else jc)
        self.gmt = dict(GMT_FIXED if gmt is None else gmt)
        d = Domain(precision, stream=stream)
        if mesh is None:
            d.AddBoxLength([0.0, 0.0, 0.0], [size, size, size], 0.5 * size / cells)
            bottom, top = workloads.box_layers(cells, cells, cells)
        else:  # (x, elnod, nodxelem, bottom, top), e.g. workloads.cylinder_tets
            x, elnod, n, bottom, top = mesh
            d.CreateFromMesh(x, elnod, n)
        self.dt = workloads.setup_compression(d, bottom, top, v_top, mat, hg_coeff=0.06)
        if cfl > 0.0:
            d.setCFL(cfl)
        self.domain, self.mat, self.nsteps, self.every = d, dict(mat), int(nsteps), int(every)
        self.x0 = d.field("x")
        self.theta = None

    def set_params(self, theta):
        a, b, hg = (float(t) for t in theta)
        m = self.mat
        d = self.domain
        if self.law == "hollomon":
            d.setMaterial(WFX_HOLLOMON, m["E"], m["nu"], m["rho"], m["sy0"], a, b)
        elif self.law == "johnson_cook":
            j = self.jc
            d.setMaterialJC(m["E"], m["nu"], m["rho"], a, b, j["n"], j["C"], j["eps_0"], j["m"], j["T_m"], j["T_t"])
            d.setTemperature(j["temp"])
        elif self.law == "gmt":
            g = self.gmt
            d.setMaterialGMT(m["E"], m["nu"], m["rho"], m["sy0"],
                             (g["n1"], b, a, g["C2"], g["m1"], g["m2"], g["I1"], g["I2"]),
                             (g["e_min"], g["e_max"], g["er_min"], g["er_max"], g["T_min"], g["T_max"]))
            d.setTemperature(g["temp"])
        else:
            d.setMaterial(WFX_BILINEAR, m["E"], m["nu"], m["rho"], a, b, 0.0)
        d.set("hg_coeff", hg)
        d.SetDT(self.dt)
        self.theta = (a, b, hg)

    def forward(self, theta):
        """reaction history for parameters theta, recorded for backward()"""
        self.set_params(theta)
        d = self.domain
        d.setField("x", self.x0)
        d.init()
        return d.adjointForward(self.nsteps, self.every)

    def backward(self, hist_bar):
        """dL/dtheta for dL/dR = hist_bar of the last forward()"""
        g = self.domain.adjointBackward(hist_bar)
        return np.array([g[k] for k in self.params])

    def loss_and_grad(self, theta, target):
        """L = sum_k (R_k - R*_k)^2 and its gradient"""
        h = self.forward(theta)
        r = h - target
        return float(np.dot(r, r)), self.backward(2.0 * r)


def _torch_function():
    import torch

    class ReactionHistory(torch.autograd.Function):
        """theta (3,) float64 tensor -> reaction history (nsteps,) tensor"""

        @staticmethod
        def forward(ctx, theta, problem):
            hist = problem.forward(theta.detach().cpu().numpy())
            ctx.problem = problem
            return torch.from_numpy(np.array(hist, dtype=np.float64)).to(theta.device)

        @staticmethod
        def backward(ctx, grad_hist):
            g = ctx.problem.backward(grad_hist.detach().cpu().numpy().astype(np.float64))
            return torch.from_numpy(g).to(grad_hist.device), None

    return ReactionHistory


_RH = None


def reaction_history(theta, problem):
    """torch entry point: problem.forward as a differentiable function of theta"""
    global _RH
    if _RH is None:
        _RH = _torch_function()
    return _RH.apply(theta, problem)


def identify(problem, target, theta0, scale=None, max_iter=100):
    """Least-squares fit of theta to a target reaction history with L-BFGS-B
    on theta/scale; returns (theta, loss, iterations, history of losses).

    The loss is divided by sum R*^2, so it falls to ~1e-10 once sigma_y0 and
    Et fit while hg_coeff, whose signal is 1e-5 of sigma_y0's in scaled
    coordinates, is still off.  The relative-reduction stop of L-BFGS-B
    compares against max(|f|, 1) and would end the fit there: it is switched
    off (ftol = gtol = 0), and the fit runs until a line search can no longer
    reduce the loss or max_iter is reached."""
    from scipy.optimize import minimize

    theta0 = np.asarray(theta0, dtype=np.float64)
    scale = np.asarray(scale if scale is not None else np.abs(theta0), dtype=np.float64)
    norm = float(np.dot(target, target))
    losses = []

    def fun(z):
        L, g = problem.loss_and_grad(z * scale, target)
        losses.append(L / norm)
        return L / norm, g * scale / norm

    res = minimize(fun, theta0 / scale, jac=True, method="L-BFGS-B",
                   bounds=[(1e-3, None)] * len(theta0), options={"maxiter": max_iter, "ftol": 0.0, "gtol": 0.0, "maxls": 50})
    return res.x * scale, float(res.fun), int(res.nit), losses
