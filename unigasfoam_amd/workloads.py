"""The five benchmark workloads: synthetic cases of the configured shapes.

heat_bath   0-D argon heat bath, 10^3 cells, specular box, 2e4 simulators
couette     1-D N2 Couette flow (LB rotational exchange), 500 cells, 1e5 simulators
cylinder    Mach-10 N2/O2 flow over a half cylinder (symmetry plane), 400x200x1
nozzle      3-D argon nozzle-type expansion, 2e6 cells, 1e8 simulators, hybrid USP/DSMC
nozzle_expansion  the same nozzle from a 1 bar reservoir towards a 1 Pa background:
            cell-weighted, started from the isentropic expansion (spec, state)
sweep       parametric Couette/Fourier cases for the batched and served forms

Each function returns (CaseSpec, fill kwargs).  Meshes are structured
planar-faced hexahedra built in polyMesh form (unigasfoam_amd.mesh).
"""
from __future__ import annotations

import math

import numpy as np

from . import abi
from .mesh import box_mesh, channel_mesh, mapped_box_mesh
from .spec import Boundary, CaseSpec, Species

KB = abi.KB

ARGON = Species("Ar", mass=6.63e-26, diameter=4.17e-10, omega=0.81, alpha=1.0, rotational_dof=0)
N2 = Species("N2", mass=4.65e-26, diameter=4.17e-10, omega=0.74, alpha=1.0, rotational_dof=2)
O2 = Species("O2", mass=5.31e-26, diameter=4.07e-10, omega=0.77, alpha=1.0, rotational_dof=2)


def heat_bath(n_cells=10, per_cell=20, n_density=1.0e20, side=0.01, dt=2.0e-6, T=(1000.0, 250.0, 250.0),
              dsmc_model="variableHardSphere", seed=1):
    """0-D heat bath: cubic box, specular walls, anisotropic Maxwellian start."""
    pt = {k: abi.WALL for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")}
    mesh = box_mesh((0.0, 0.0, 0.0), (side, side, side), (n_cells, n_cells, n_cells), pt)
    n_sim = per_cell * n_cells ** 3
    fn = n_density * side ** 3 / n_sim
    bnd = [Boundary(p, abi.BND_SPECULAR) for p in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")]
    spec = CaseSpec(mesh=mesh, species=[ARGON], delta_t=dt, n_equivalent_particles=fn,
                    collision_model="dsmc", dsmc_model=dsmc_model, Tref=273.0, boundaries=bnd, seed=seed)
    fill = dict(number_density={"Ar": n_density}, T=T, exact_per_cell=per_cell)
    return spec, fill


def couette(n_cells=500, per_cell=200, gap=1.0e-3, n_density=1.3e22, T_wall=273.0, U_wall=250.0,
            dt=5.0e-9, gas="N2", dsmc_model="LarsenBorgnakkeVariableHardSphere", periodic="cyclic", seed=2,
            width=None, collision_model="dsmc", bgk_model="noBGKCollision", T_lower=None, T_upper=None):
    """1-D Couette flow between diffuse walls moving at -U_wall (y=0) and +U_wall (y=gap)."""
    width = width if width is not None else gap / n_cells
    mesh = channel_mesh(n_cells, gap, width, periodic=periodic)
    sp = N2 if gas == "N2" else ARGON
    if gas != "N2" and dsmc_model.startswith("LarsenBorgnakke"):
        dsmc_model = "variableHardSphere"
    vol = width * gap * width
    fn = n_density * vol / (per_cell * n_cells)
    Tl = T_wall if T_lower is None else T_lower
    Tu = T_wall if T_upper is None else T_upper
    bnd = [Boundary("lowerWall", abi.BND_DIFFUSE, temperature=Tl, velocity=(-U_wall, 0.0, 0.0)),
           Boundary("upperWall", abi.BND_DIFFUSE, temperature=Tu, velocity=(U_wall, 0.0, 0.0))]
    spec = CaseSpec(mesh=mesh, species=[sp], delta_t=dt, n_equivalent_particles=fn,
                    collision_model=collision_model, dsmc_model=dsmc_model, bgk_model=bgk_model,
                    Tref=273.0, Zrot=5.0, boundaries=bnd, seed=seed)
    fill = dict(number_density={sp.name: n_density}, T=0.5 * (Tl + Tu), exact_per_cell=per_cell)
    return spec, fill


def half_annulus_mesh(n_theta, n_r, r0, R, depth, grading=10.0):
    """Half O-grid around a cylinder (y >= 0), extruded one cell in z."""
    m = n_r
    if grading != 1.0 and m > 1:
        rr = grading ** (1.0 / (m - 1))
        w = rr ** np.arange(m)
    else:
        w = np.ones(m)
    e = np.concatenate([[0.0], np.cumsum(w)]) / np.sum(w)

    def fn(S):
        s, t, u = S[..., 0], S[..., 1], S[..., 2]
        ti = np.rint(t * m).astype(int)
        r = r0 + (R - r0) * e[ti]
        th = math.pi * (1.0 - s)
        return np.stack([r * np.cos(th), r * np.sin(th), depth * u], axis=-1)

    pt = {"xmin": abi.SYMMETRY, "xmax": abi.SYMMETRY, "ymin": abi.WALL, "ymax": abi.PATCH,
          "zmin": abi.EMPTY, "zmax": abi.EMPTY}
    names = {"xmin": "axisLeft", "xmax": "axisRight", "ymin": "cylinder", "ymax": "freestream",
             "zmin": "front", "zmax": "back"}
    return mapped_box_mesh((n_theta, n_r, 1), fn, pt, names)


def cylinder(n_theta=400, n_r=200, r0=0.1, R=1.0, depth=0.01, n_inf=1.0e21, T_inf=200.0, Mach=10.0,
             T_wall=500.0, target=5.0e7, dt=5.0e-8, seed=3):
    """Mach-10 N2/O2 (79/21) over a cylinder; diffuse 500 K wall; freestream inflow + deletion outside."""
    mesh = half_annulus_mesh(n_theta, n_r, r0, R, depth)
    xs = {"N2": 0.79, "O2": 0.21}
    m_mix = xs["N2"] * N2.mass + xs["O2"] * O2.mass
    a = math.sqrt(1.4 * KB / m_mix * T_inf)
    U = Mach * a
    vol = 0.5 * math.pi * (R * R - r0 * r0) * depth
    fn = n_inf * vol / target
    nd = {k: n_inf * v for k, v in xs.items()}
    bnd = [Boundary("cylinder", abi.BND_DIFFUSE, temperature=T_wall),
           Boundary("freestream", abi.BND_DELETION),
           Boundary("freestream", abi.BND_FREESTREAM, temperature=T_inf, rot_temperature=T_inf,
                    velocity=(U, 0.0, 0.0), number_density=nd)]
    spec = CaseSpec(mesh=mesh, species=[N2, O2], delta_t=dt, n_equivalent_particles=fn,
                    collision_model="dsmc", dsmc_model="LarsenBorgnakkeVariableHardSphere", Tref=273.0,
                    Zrot=5.0, boundaries=bnd, seed=seed)
    fill = dict(number_density=nd, T=T_inf, rot_T=T_inf, velocity=(U, 0.0, 0.0))
    return spec, fill


def nozzle_half_width(x, L=0.02, h_in=4.0e-3, h_throat=1.0e-3, h_out=6.0e-3):
    """Half-width h(x) of the nozzle's square cross-section: h_in at the inlet,
    h_throat at the throat x = L/4, h_out at the exit."""
    xt = 0.25 * L
    return np.where(x < xt, h_throat + (h_in - h_throat) * ((xt - x) / xt) ** 2,
                    h_throat + (h_out - h_throat) * np.clip((x - xt) / (L - xt), 0.0, None) ** 1.5)


def nozzle_mesh(nx, ny, nz, L=0.02, h_in=4.0e-3, h_throat=1.0e-3, h_out=6.0e-3):
    """Converging-diverging channel: square cross-section of half-width h(x),
    throat at x = L/4; cross-sections scaled along x keep every face planar."""
    def h(x):
        return nozzle_half_width(x, L, h_in, h_throat, h_out)

    def fn(S):
        s, t, u = S[..., 0], S[..., 1], S[..., 2]
        x = L * s
        hx = h(x)
        return np.stack([x, hx * (2.0 * t - 1.0), hx * (2.0 * u - 1.0)], axis=-1)

    pt = {"xmin": abi.PATCH, "xmax": abi.PATCH, "ymin": abi.WALL, "ymax": abi.WALL,
          "zmin": abi.WALL, "zmax": abi.WALL}
    names = {"xmin": "inlet", "xmax": "outlet"}
    return mapped_box_mesh((nx, ny, nz), fn, pt, names)


def nozzle(nx=200, ny=100, nz=100, per_cell=50, n_stag=2.4e25, T0=300.0, n_back=2.4e20, dt=1.0e-8,
           collision_model="hybrid", bgk_model="unifiedStochasticParticleSBGK", seed=4):
    """Argon nozzle expansion with per-cell DSMC/USP switching.  Uniform
    simulator weight (the case does not turn on cell weighting), so the fill
    uses one stagnation-side density everywhere and the inlet keeps feeding it."""
    mesh = nozzle_mesh(nx, ny, nz)
    n_cells = nx * ny * nz
    # volume of the channel
    xs = np.linspace(0.0, 0.02, 2001)
    xt = 0.005
    hh = np.where(xs < xt, 1e-3 + 3e-3 * ((xt - xs) / xt) ** 2, 1e-3 + 5e-3 * np.clip((xs - xt) / 0.015, 0.0, None) ** 1.5)
    vol = float(np.trapezoid((2 * hh) ** 2, xs))
    n_fill = n_stag * 1.0e-3       # expanded state filling the channel initially
    fn = n_fill * vol / (per_cell * n_cells)
    bnd = [Boundary("inlet", abi.BND_DELETION),
           Boundary("inlet", abi.BND_FREESTREAM, temperature=T0, velocity=(50.0, 0.0, 0.0),
                    number_density={"Ar": n_fill}),
           Boundary("outlet", abi.BND_DELETION),
           Boundary("ymin", abi.BND_DIFFUSE, temperature=T0), Boundary("ymax", abi.BND_DIFFUSE, temperature=T0),
           Boundary("zmin", abi.BND_DIFFUSE, temperature=T0), Boundary("zmax", abi.BND_DIFFUSE, temperature=T0)]
    spec = CaseSpec(mesh=mesh, species=[ARGON], delta_t=dt, n_equivalent_particles=fn,
                    collision_model=collision_model, dsmc_model="variableHardSphere",
                    bgk_model=bgk_model if collision_model in ("bgk", "hybrid") else "noBGKCollision",
                    Tref=273.0, boundaries=bnd, seed=seed, decomposition_interval=20, smoothing_passes=5,
                    breakdown_max=0.05, kn_theta=0.5)
    fill = dict(number_density={"Ar": n_fill}, T=T0, exact_per_cell=per_cell)
    return spec, fill


def isentropic_mach(area_ratio, supersonic):
    """Mach number of isentropic quasi-one-dimensional flow of a monatomic
    gas (gamma = 5/3) at area ratio A/A* >= 1.  For gamma = 5/3 the
    area-Mach relation (1/M) [(2/(g+1))(1 + (g-1)/2 M^2)]^((g+1)/(2(g-1)))
    is ((3 + M^2)/4)^2 / M; it is bisected on the subsonic or the supersonic
    branch."""
    r = np.asarray(area_ratio, dtype=np.float64)
    sup = np.broadcast_to(np.asarray(supersonic, dtype=bool), r.shape)
    lo = np.where(sup, 1.0, 1.0e-9)
    hi = np.where(sup, 100.0, 1.0)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        f = ((3.0 + mid * mid) / 4.0) ** 2 / mid - r
        up = np.where(sup, f < 0.0, f > 0.0)      # the root lies above mid
        lo = np.where(up, mid, lo)
        hi = np.where(up, hi, mid)
    return 0.5 * (lo + hi)


def nozzle_expansion(nx=200, ny=100, nz=100, per_cell=50, p0=1.0e5, T0=300.0, p_back=1.0, T_back=300.0,
                     dt=1.0e-8, bgk_model="unifiedStochasticParticleSBGK", seed=4, L=0.02, h_in=4.0e-3,
                     h_throat=1.0e-3, h_out=6.0e-3):
    """Argon expanding from a 1 bar, 300 K reservoir through the nozzle's
    2 mm throat towards a 1 Pa background, with cell weighting so that every
    cell holds per_cell simulators although the density falls by two orders
    of magnitude along the nozzle (and by five to the background).

    The initial state is the isentropic quasi-one-dimensional expansion:
    per cell column, the Mach number of the column's area ratio (subsonic
    before the throat, supersonic after it), T = T0 / (1 + M^2 / 3),
    n = n0 (T / T0)^(3/2) and U = M sqrt(5 k T / (3 m)) along x.  The inlet
    feeds the reservoir gas at the inlet state, the outlet the 1 Pa
    background; both also delete what leaves.  Returns (spec, state): the
    state's cell weights, (sigma_T c_r)max and simulators are applied to a
    case by apply_state."""
    mesh = nozzle_mesh(nx, ny, nz, L, h_in, h_throat, h_out)
    m = ARGON.mass
    n0 = p0 / (KB * T0)
    xs = L * np.arange(nx + 1) / nx                       # column faces
    hf = nozzle_half_width(xs, L, h_in, h_throat, h_out)
    xc = 0.5 * (xs[:-1] + xs[1:])
    hc = 0.5 * (hf[:-1] + hf[1:])
    Mc = isentropic_mach((hc / h_throat) ** 2, xc > 0.25 * L)
    Tc = T0 / (1.0 + Mc * Mc / 3.0)
    nc_col = n0 * (Tc / T0) ** 1.5
    Uc = Mc * np.sqrt(5.0 * KB * Tc / (3.0 * m))
    # column volumes: the cross-section's half-width is linear in x between
    # the column's faces, so V = (4 / (ny nz)) dx (h0^2 + h0 h1 + h1^2) / 3
    dx = L / nx
    vol_col = 4.0 / (ny * nz) * dx * (hf[:-1] ** 2 + hf[:-1] * hf[1:] + hf[1:] ** 2) / 3.0
    # F_N so that an inlet cell has weight 1
    fn = nc_col[0] * vol_col[0] / per_cell
    wcol = nc_col * vol_col / (per_cell * fn)
    # inlet state: the first column's (the reservoir gas flowing in)
    M_in = float(isentropic_mach((h_in / h_throat) ** 2, False))
    T_in = T0 / (1.0 + M_in * M_in / 3.0)
    n_in = n0 * (T_in / T0) ** 1.5
    U_in = M_in * math.sqrt(5.0 * KB * T_in / (3.0 * m))
    n_back = p_back / (KB * T_back)
    bnd = [Boundary("inlet", abi.BND_DELETION),
           Boundary("inlet", abi.BND_FREESTREAM, temperature=T_in, velocity=(U_in, 0.0, 0.0),
                    number_density={"Ar": n_in}),
           Boundary("outlet", abi.BND_DELETION),
           Boundary("outlet", abi.BND_FREESTREAM, temperature=T_back, velocity=(0.0, 0.0, 0.0),
                    number_density={"Ar": n_back}),
           Boundary("ymin", abi.BND_DIFFUSE, temperature=T0), Boundary("ymax", abi.BND_DIFFUSE, temperature=T0),
           Boundary("zmin", abi.BND_DIFFUSE, temperature=T0), Boundary("zmax", abi.BND_DIFFUSE, temperature=T0)]
    spec = CaseSpec(mesh=mesh, species=[ARGON], delta_t=dt, n_equivalent_particles=fn,
                    collision_model="hybrid", dsmc_model="variableHardSphere", bgk_model=bgk_model,
                    Tref=273.0, boundaries=bnd, seed=seed, decomposition_interval=20, smoothing_passes=5,
                    breakdown_max=0.05, kn_theta=0.5, cell_weighted=True, particles_per_subcell=per_cell)
    ncell = nx * ny * nz
    col = np.arange(ncell) % nx
    # simulators: per_cell in every cell, uniform in the cell's parameters
    # (the trilinear map of its corners), Maxwellian at the column's state
    rng = np.random.default_rng(seed)
    cell = np.repeat(np.arange(ncell, dtype=np.int32), per_cell)
    i = cell % nx
    j = (cell // nx) % ny
    k = cell // (nx * ny)
    a = rng.random(len(cell))
    x = xs[i] + a * dx
    hl = hf[i] + a * (hf[i + 1] - hf[i])
    y = hl * (2.0 * (j + rng.random(len(cell))) / ny - 1.0)
    z = hl * (2.0 * (k + rng.random(len(cell))) / nz - 1.0)
    del a
    pos = np.stack([x, y, z], axis=1)
    del x, y, z, hl, j, k
    sd = np.sqrt(KB * Tc[i] / m)
    vel = rng.standard_normal((len(cell), 3)) * sd[:, None]
    vel[:, 0] += Uc[i]
    del sd, i
    d = ARGON.diameter
    state = {"cell_weights": wcol[col], "sigma_max": math.pi * d * d * np.sqrt(2.0 * KB * Tc[col] / m),
             "position": pos, "velocity": vel, "cell": cell,
             "columns": {"x": xc, "M": Mc, "T": Tc, "n": nc_col, "U": Uc, "volume": vol_col, "weight": wcol}}
    return spec, state


def apply_state(case, state):
    """Give a new case (GPU or oracle) the cell weights, (sigma_T c_r)max and
    simulators of a nozzle_expansion state."""
    case.set_cell_weights(state["cell_weights"])
    case.set_sigma_max(state["sigma_max"])
    case.upload(velocity=state["velocity"], cell=state["cell"], position=state["position"])


def sweep_params(n_cases=256, seed=5):
    """Parametric Couette/Fourier draws: Kn log-uniform [0.01, 10], wall T
    uniform [200, 1000] K, wall speed uniform [0, 500] m/s, cells [50, 2000],
    simulators/cell [20, 200], argon or N2 with equal probability."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_cases):
        out.append(dict(
            Kn=float(10 ** rng.uniform(-2, 1)),
            T_wall=float(rng.uniform(200, 1000)),
            U_wall=float(rng.uniform(0, 500)),
            n_cells=int(rng.integers(50, 2001)),
            per_cell=int(rng.integers(20, 201)),
            gas="Ar" if rng.random() < 0.5 else "N2",
            seed=1000 + i,
        ))
    return out


def sweep_case(p, gap=1.0e-3):
    """One parametric case: number density from Kn = lambda/gap (VHS mean free
    path at the wall temperature)."""
    sp = ARGON if p["gas"] == "Ar" else N2
    lam = p["Kn"] * gap
    T = p["T_wall"]
    om = sp.omega
    # Bird eq. 4.65 for the VHS mean free path at T with Tref = 273 K
    n = 1.0 / (math.sqrt(2.0) * math.pi * sp.diameter ** 2 * lam * (273.0 / T) ** (om - 0.5))
    dx = gap / p["n_cells"]
    c_mp = math.sqrt(2 * KB * T / sp.mass)
    dt = 0.3 * min(dx, lam) / (c_mp + p["U_wall"])
    return couette(n_cells=p["n_cells"], per_cell=p["per_cell"], gap=gap, n_density=n, T_wall=T,
                   U_wall=0.5 * p["U_wall"], dt=dt, gas=p["gas"], seed=p["seed"])
