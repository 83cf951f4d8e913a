"""Seeded synthetic NS3D inputs and HyPar input decks.

The reference never defines its benchmark fields, so this module fixes them
(numpy PCG64 ``default_rng(seed)``, draw order field -> mode -> (k, phi)) and
writes them in the reference's own input formats so that the same deck runs
on the reference CPU binary, on this package's C driver and through the
Python layer:

* ``solver.inp``   keywords of src/Simulation/ReadInputs.c:152-430
* ``physics.inp``  keywords of src/PhysicalModels/NavierStokes3D/NavierStokes3DInitialize.c:162-333
* ``weno.inp``     keywords of src/InterpolationFunctions/WENOInitialize.c:62-96
* ``boundary.inp`` six ``periodic`` faces (src/BoundaryConditions/BCInitialize.c)
* ``initial.inp``  binary: x, y, z then u with the 5 variables fastest, then
  x, y, z (src/IOFunctions/ReadArray.c:215-256)
* ``op.bin``       binary output (src/IOFunctions/WriteBinary.c:34-100)

Grid metrics follow the reference exactly: ghost coordinates are linearly
extrapolated (src/IOFunctions/ReadArray.c:59-90), interior
``dxinv[i] = 2/(x[i+1]-x[i-1])`` and ghost ``dxinv`` copies the nearest
interior value (src/Simulation/InitialSolution.c:74-113).
"""
from __future__ import annotations

import dataclasses
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

GAMMA = 1.4
NVARS = 5
GHOSTS = 3
TWO_PI = 2.0 * np.pi


# --------------------------------------------------------------------------
# workloads
# --------------------------------------------------------------------------
@dataclasses.dataclass
class Workload:
    """One single-grid NS3D run: grid, numerics and physics of a deck."""

    name: str
    dims: Tuple[int, int, int]
    interp: str = "components"        # hyp_interp_type: components | characteristic
    upwind: str = "rusanov"           # physics.inp upwinding: rusanov | roe
    Re: float = -1.0                  # physics.inp Re (scaled by 1/Minf inside the solver)
    Minf: float = 1.0
    Pr: float = 0.72
    gamma: float = GAMMA
    weno_eps: float = 1e-6
    weno_type: str = "js"             # js | mapped | borges | yc | optimal (no_limiting 1)
    n_iter: int = 100
    cfl: float = 0.5
    seed: int = 20260
    field: str = "random_phase"       # random_phase | taylor_green
    dt: Optional[float] = None        # filled from the CFL when None
    time_scheme_type: str = "44"      # solver.inp time_scheme_type (rk class)

    @property
    def viscous(self) -> bool:
        return self.Re > 0

    @property
    def ncells(self) -> int:
        return int(np.prod(self.dims))


def config0(N: int = 64, n_iter: int = 100) -> Workload:
    """Euler, component-wise WENO5-JS + Rusanov, RK4, CFL 0.5 (seed 20260)."""
    return Workload("euler-weno5js-components-rusanov", (N, N, N), n_iter=n_iter,
                    seed=20260, field="random_phase")


def config1(N: int = 256, n_iter: int = 200) -> Workload:
    """Viscous NS3D (Re 1600, Pr 0.72, Mach 0.1), characteristic WENO5-JS + Roe,
    4th-order viscous terms, RK4, CFL 0.5 (seed 20261)."""
    return Workload("ns3d-weno5js-char-roe-visc", (N, N, N), interp="characteristic",
                    upwind="roe", Re=1600.0, Minf=0.1, Pr=0.72, n_iter=n_iter,
                    seed=20261, field="taylor_green")


def config_dns(N: int = 256, n_iter: int = 1000) -> Workload:
    """The scheme of the reference's published GPU example deck
    (Examples/3D/NavierStokes3D/DNS_IsotropicTurbulenceDecay_CUDA:
    solver.inp:9-15, weno.inp:2, physics.inp:2-6): viscous NS3D (Re 333.33,
    Mach 0.3, Pr 0.72), component-wise mapped WENO5 + Roe, SSPRK3, dt 0.005.
    The deck's isotropic-turbulence initial condition needs FFTW
    (Examples/.../aux/init.c), so the field is the seeded Taylor-Green type
    at Mach 0.3 (seed 20265)."""
    return Workload("dns-weno5mapped-components-roe-visc-ssprk3", (N, N, N), interp="components",
                    upwind="roe", Re=333.333333333333333, Minf=0.3, Pr=0.72, weno_type="mapped",
                    n_iter=n_iter, seed=20265, field="taylor_green", dt=0.005, time_scheme_type="ssprk3")


def config3(N: int = 128, n_iter: int = 50) -> Workload:
    """Adjoint workload: config-0 physics at 128^3 (seed 20263)."""
    return Workload("euler-adjoint", (N, N, N), n_iter=n_iter, seed=20263,
                    field="random_phase")


def ensemble_config2(n_members: int = 64, sides: Sequence[int] = (32, 48, 64, 96, 128),
                     n_iter: int = 100, seed: int = 20262) -> List[Workload]:
    """64 Euler members with uneven periodic grids, each with its own dt."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n_members):
        d = tuple(int(sides[i]) for i in rng.integers(0, len(sides), size=3))
        out.append(Workload(f"member{m}", d, n_iter=n_iter, seed=seed * 1000 + m,
                            field="random_phase"))
    return out


def ensemble_config4(n_members: int = 32, sides: Sequence[int] = (64, 96, 128),
                     n_iter: int = 30, seed: int = 20264) -> List[Workload]:
    """Viscous (Re 1000, Mach 0.1) characteristic WENO5 + Roe members."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n_members):
        d = tuple(int(sides[i]) for i in rng.integers(0, len(sides), size=3))
        out.append(Workload(f"member{m}", d, interp="characteristic", upwind="roe",
                            Re=1000.0, Minf=0.1, n_iter=n_iter, seed=seed * 1000 + m,
                            field="taylor_green"))
    return out


# --------------------------------------------------------------------------
# fields
# --------------------------------------------------------------------------
_MODE_TABLES = {}


def mode_table(K: int) -> np.ndarray:
    """All integer wave vectors k with 0 < |k|_2 <= K, in a fixed order."""
    if K not in _MODE_TABLES:
        r = range(-K, K + 1)
        ks = [(a, b, c) for a in r for b in r for c in r if 0 < a * a + b * b + c * c <= K * K]
        _MODE_TABLES[K] = np.array(ks, dtype=np.int64)
    return _MODE_TABLES[K]


def draw_modes(rng: np.random.Generator, K: int, nmodes: int):
    table = mode_table(K)
    modes = []
    for _ in range(nmodes):
        k = table[int(rng.integers(0, len(table)))]
        phi = float(rng.uniform(0.0, TWO_PI))
        modes.append((k, phi))
    return modes


def grid_1d(n: int) -> np.ndarray:
    """Periodic grid on [0, 2pi): x_i = i * (2pi/n)."""
    return np.arange(n, dtype=np.float64) * (TWO_PI / n)


def _mesh(dims):
    x, y, z = (grid_1d(n) for n in dims)
    # arrays indexed [k][j][i] (x fastest in memory)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def _sum_modes(modes, X, Y, Z):
    s = np.zeros_like(X)
    for k, phi in modes:
        s += np.sin(k[0] * X + k[1] * Y + k[2] * Z + phi)
    return s


def primitive_fields(w: Workload):
    """rho, u, v, w, p on the interior grid, arrays indexed [k][j][i]."""
    rng = np.random.default_rng(w.seed)
    X, Y, Z = _mesh(w.dims)
    g = w.gamma
    if w.field == "random_phase":
        mrho = draw_modes(rng, 3, 4)
        mu = draw_modes(rng, 3, 4)
        mv = draw_modes(rng, 3, 4)
        mw = draw_modes(rng, 3, 4)
        rho = 1.0 + 0.1 * _sum_modes(mrho, X, Y, Z)
        u = 0.2 * _sum_modes(mu, X, Y, Z)
        v = 0.2 * _sum_modes(mv, X, Y, Z)
        ww = 0.2 * _sum_modes(mw, X, Y, Z)
        p = np.full_like(X, 1.0 / g)
    elif w.field == "taylor_green":
        M = w.Minf
        mu = draw_modes(rng, 8, 4)
        mv = draw_modes(rng, 8, 4)
        mw = draw_modes(rng, 8, 4)
        amp = 0.05 * M
        u = M * np.sin(X) * np.cos(Y) * np.cos(Z) + amp * _sum_modes(mu, X, Y, Z)
        v = -M * np.cos(X) * np.sin(Y) * np.cos(Z) + amp * _sum_modes(mv, X, Y, Z)
        ww = amp * _sum_modes(mw, X, Y, Z)
        p = 1.0 / g + (M * M / 16.0) * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
        rho = np.ones_like(X)
    else:
        raise ValueError(f"unknown field {w.field!r}")
    return rho, u, v, ww, p


def conserved_from_primitive(rho, u, v, w, p, gamma=GAMMA) -> np.ndarray:
    """Stack to conserved variables, shape (nz, ny, nx, 5) (AoS, x fastest)."""
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.ascontiguousarray(np.stack([rho, rho * u, rho * v, rho * w, E], axis=-1))


def initial_state(w: Workload) -> np.ndarray:
    """Conserved initial state of a workload, shape (nz, ny, nx, 5), float64."""
    return conserved_from_primitive(*primitive_fields(w), gamma=w.gamma)


def density_perturbation(dims, seed: int, amplitude: float = 1e-2, K: int = 4,
                         nmodes: int = 4) -> np.ndarray:
    """Seeded initial density perturbation field (adjoint control), [k][j][i]."""
    rng = np.random.default_rng(seed)
    X, Y, Z = _mesh(dims)
    return amplitude * _sum_modes(draw_modes(rng, K, nmodes), X, Y, Z)


# --------------------------------------------------------------------------
# grid metrics (exactly as the reference computes them)
# --------------------------------------------------------------------------
def ghosted_coords(x: np.ndarray, ghosts: int = GHOSTS) -> np.ndarray:
    """Linear extrapolation of coordinates into the ghost layers
    (src/IOFunctions/ReadArray.c:59-90)."""
    n = len(x)
    X = np.zeros(n + 2 * ghosts)
    X[ghosts:ghosts + n] = x
    for i in range(ghosts):
        delta = ghosts - i
        X[i] = X[ghosts] + float(delta) * (X[ghosts] - X[ghosts + 1])
    for i in range(n + ghosts, n + 2 * ghosts):
        delta = i - (n + ghosts - 1)
        X[i] = X[n + ghosts - 1] + float(delta) * (X[n + ghosts - 1] - X[n + ghosts - 2])
    return X


def dxinv_1d(x: np.ndarray, ghosts: int = GHOSTS) -> np.ndarray:
    """Per-point inverse spacing with ghosts (src/Simulation/InitialSolution.c:74-113)."""
    X = ghosted_coords(x, ghosts)
    n = len(x)
    d = np.zeros(n + 2 * ghosts)
    for i in range(n):
        d[i + ghosts] = 2.0 / (X[i + 1 + ghosts] - X[i - 1 + ghosts])
    d[:ghosts] = d[ghosts]
    d[n + ghosts:] = d[n + ghosts - 1]
    return d


def dxinv_all(dims, ghosts: int = GHOSTS) -> np.ndarray:
    """Concatenated dxinv (x, then y, then z), each with ghosts, as HyPar stores it."""
    return np.concatenate([dxinv_1d(grid_1d(n), ghosts) for n in dims])


# --------------------------------------------------------------------------
# time step from the CFL condition
# --------------------------------------------------------------------------
def cfl_dt(u: np.ndarray, dims, gamma: float = GAMMA, cfl: float = 0.5) -> float:
    """dt with max_cell,dir (|v_d|+c) dt dxinv_d = cfl
    (src/PhysicalModels/NavierStokes3D/NavierStokes3DComputeCFL.c:39-47)."""
    rho = u[..., 0]
    vx, vy, vz = u[..., 1] / rho, u[..., 2] / rho, u[..., 3] / rho
    P = (u[..., 4] - 0.5 * rho * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0)
    c = np.sqrt(gamma * P / rho)
    d = [dxinv_1d(grid_1d(n))[GHOSTS:GHOSTS + n] for n in dims]
    m = max(float(np.max((np.abs(vx) + c) * d[0][None, None, :])),
            float(np.max((np.abs(vy) + c) * d[1][None, :, None])),
            float(np.max((np.abs(vz) + c) * d[2][:, None, None])))
    return float(cfl / m)


def prepare(w: Workload):
    """Initial state and dt of a workload (dt from the CFL when unset)."""
    u0 = initial_state(w)
    if w.dt is None:
        w.dt = cfl_dt(u0, w.dims, w.gamma, w.cfl)
    return u0, w.dt


# --------------------------------------------------------------------------
# deck files
# --------------------------------------------------------------------------
def _weno_flags(weno_type: str):
    """(mapped, borges, yc, no_limiting) of weno.inp"""
    return {"js": (0, 0, 0, 0), "mapped": (1, 0, 0, 0), "borges": (0, 1, 0, 0),
            "yc": (0, 0, 1, 0), "optimal": (0, 0, 0, 1)}[weno_type]


def write_solver_inp(path: str, dims_list, n_iter: int, dt: float, interp: str,
                     screen_op_iter: int = 1, file_op_iter: Optional[int] = None,
                     time_scheme_type: str = "44", conservation_check: str = "no"):
    size = " ".join(" ".join(str(n) for n in d) for d in dims_list)
    with open(os.path.join(path, "solver.inp"), "w") as f:
        f.write("begin\n")
        f.write("  ndims               3\n")
        f.write("  nvars               5\n")
        f.write(f"  size                {size}\n")
        f.write("  iproc               " + " ".join(["1 1 1"] * len(dims_list)) + "\n")
        f.write("  ghost               3\n")
        f.write(f"  n_iter              {n_iter}\n")
        f.write("  restart_iter        0\n")
        f.write("  time_scheme         rk\n")
        f.write(f"  time_scheme_type    {time_scheme_type}\n")
        f.write("  hyp_space_scheme    weno5\n")
        f.write(f"  hyp_interp_type     {interp}\n")
        f.write("  par_space_type      nonconservative-2stage\n")
        f.write("  par_space_scheme    4\n")
        f.write(f"  dt                  {dt:.17g}\n")
        f.write(f"  conservation_check  {conservation_check}\n")
        f.write(f"  screen_op_iter      {screen_op_iter}\n")
        f.write(f"  file_op_iter        {file_op_iter if file_op_iter else n_iter}\n")
        f.write("  ip_file_type        binary\n")
        f.write("  input_mode          serial\n")
        f.write("  output_mode         serial\n")
        f.write("  op_file_format      binary\n")
        f.write("  op_overwrite        yes\n")
        f.write("  model               navierstokes3d\n")
        f.write("end\n")


def write_physics_inp(path: str, w: Workload):
    with open(os.path.join(path, "physics.inp"), "w") as f:
        f.write("begin\n")
        f.write(f"  gamma     {w.gamma:.17g}\n")
        f.write(f"  upwinding {w.upwind}\n")
        f.write(f"  Pr        {w.Pr:.17g}\n")
        f.write(f"  Minf      {w.Minf:.17g}\n")
        f.write(f"  Re        {w.Re:.17g}\n")
        f.write("end\n")


def write_weno_inp(path: str, w: Workload):
    m, b, y, nl = _weno_flags(w.weno_type)
    with open(os.path.join(path, "weno.inp"), "w") as f:
        f.write("begin\n")
        f.write(f"  mapped        {m}\n")
        f.write(f"  borges        {b}\n")
        f.write(f"  yc            {y}\n")
        f.write(f"  no_limiting   {nl}\n")
        f.write(f"  epsilon       {w.weno_eps:.17g}\n")
        f.write("  p             2.0\n")
        f.write("end\n")


def write_boundary_inp(path: str, fname: str = "boundary.inp"):
    L = TWO_PI
    with open(os.path.join(path, fname), "w") as f:
        f.write("6\n")
        f.write(f"periodic 0  1 0 0   0 {L} 0 {L}\n")
        f.write(f"periodic 0 -1 0 0   0 {L} 0 {L}\n")
        f.write(f"periodic 1  1 0 {L} 0 0   0 {L}\n")
        f.write(f"periodic 1 -1 0 {L} 0 0   0 {L}\n")
        f.write(f"periodic 2  1 0 {L} 0 {L} 0 0\n")
        f.write(f"periodic 2 -1 0 {L} 0 {L} 0 0\n")


def write_initial(path: str, dims, u: np.ndarray, fname: str = "initial.inp"):
    """Binary initial solution: x, y, z, then u (AoS, x fastest)."""
    assert u.shape == (dims[2], dims[1], dims[0], NVARS)
    with open(os.path.join(path, fname), "wb") as f:
        for n in dims:
            grid_1d(n).astype("<f8").tofile(f)
        np.ascontiguousarray(u, dtype="<f8").tofile(f)


def write_deck(path: str, w: Workload, u0: Optional[np.ndarray] = None,
               screen_op_iter: int = 1, conservation_check: str = "no") -> np.ndarray:
    """Write a complete single-simulation deck; returns the initial state."""
    os.makedirs(path, exist_ok=True)
    if u0 is None:
        u0, _ = prepare(w)
    elif w.dt is None:
        w.dt = cfl_dt(u0, w.dims, w.gamma, w.cfl)
    write_solver_inp(path, [w.dims], w.n_iter, w.dt, w.interp, screen_op_iter,
                     time_scheme_type=w.time_scheme_type, conservation_check=conservation_check)
    write_physics_inp(path, w)
    write_weno_inp(path, w)
    write_boundary_inp(path)
    write_initial(path, w.dims, u0)
    return u0


def write_ensemble_deck(path: str, members: Sequence[Workload], n_iter: int, dt: float):
    """Ensemble deck (simulation.inp with nsims, per-member initial_<i>.inp;
    shared dt and n_iter as in src/TimeIntegration/TimeInitialize.c:49-52)."""
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "simulation.inp"), "w") as f:
        f.write(f"begin\n  nsims {len(members)}\nend\n")
    write_solver_inp(path, [m.dims for m in members], n_iter, dt, members[0].interp)
    write_physics_inp(path, members[0])
    write_weno_inp(path, members[0])
    write_boundary_inp(path)
    for i, m in enumerate(members):
        write_initial(path, m.dims, initial_state(m), f"initial_{member_index(i, len(members))}.inp")


def member_index(i: int, nsims: int) -> str:
    """Zero-padded member index used in ensemble file names
    (GetStringFromInteger with width int(log10(nsims))+1)."""
    return str(i).zfill(int(np.log10(nsims)) + 1)


def read_op_bin(fname: str):
    """Read a binary solution file written by WriteBinary; returns (dims, x, u)
    with u shaped (nz, ny, nx, nvars)."""
    with open(fname, "rb") as f:
        ndims, nvars = np.fromfile(f, dtype="<i4", count=2)
        dims = tuple(int(v) for v in np.fromfile(f, dtype="<i4", count=ndims))
        x = np.fromfile(f, dtype="<f8", count=sum(dims))
        u = np.fromfile(f, dtype="<f8", count=int(np.prod(dims)) * nvars)
    return dims, x, u.reshape(dims[2], dims[1], dims[0], nvars)


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    """Relative L2 error ||a-b|| / ||b|| over all entries."""
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
