"""Synthetic, seeded workloads of the benchmark configurations.

Random numbers come from a counter-based SplitMix64 (Steele, Lea, Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014): the value for
(seed, stream, index) is mix(seed + (stream * 2^40 + index + 1) * 0x9E3779B97F4A7C15),
so any implementation reproduces the same mesh and coefficients.

* heat_t3:      unit square, N x N T3, kappa = 1, Q = -6, Dirichlet
                u = 1 + x^2 + 2 y^2 on the boundary (test/test_heat.jl:22-26).
* elast_t6:     unit square, N x N T6, 2 dofs/node, plane strain E = 1,
                nu = 0.3; every node jittered by h * U(-j, j) per axis, boundary
                nodes only tangentially; x = 0 clamped, x = 1 displaced
                (the boundary conditions of test/test_elasticity.jl:191-204).
* stokes_t6t3:  unit square, N x N Taylor-Hood T6/T3, mu = 1, all boundary
                velocity dofs prescribed, pressure pinned at the node nearest
                the centre (test/test_stokes.jl:740-751).
* heat_q4:      N x N Q4, kappa_e ~ lognormal(0, sigma), f_e ~ U[-1, 1], u = 0
                on the boundary.
* batch:        many heat-T3 / elasticity-T6 problems of uneven size.
"""
import numpy as np

from .felements import FEH1_Q4, FEH1_T3, FEH1_T6
from .fespaces import FESpace, numberdofs
from .integrate import plane_strain_D
from .meshes import Q4block, T3block, T6block, T6toT3

SEED = 20261202
_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(x):
    x = np.asarray(x, np.uint64)
    with np.errstate(over="ignore"):
        z = x + _GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, stream, index):
    """U[0, 1) for each index of a stream."""
    idx = np.asarray(index, np.uint64)
    with np.errstate(over="ignore"):
        ctr = np.uint64(seed) + (np.uint64(stream) * np.uint64(1 << 40) + idx + np.uint64(1)) * _GAMMA
    return (splitmix64(ctr) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, stream, n):
    """Box-Muller standard normals from uniforms 2i, 2i+1 of the stream."""
    i = np.arange(n, dtype=np.uint64)
    u1 = uniform(seed, stream, 2 * i)
    u2 = uniform(seed, stream, 2 * i + 1)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2 * np.pi * u2)


def jitter(mesh, h, frac, seed=SEED, stream=1, box=(0.0, 1.0, 0.0, 1.0)):
    """x += h * U(-frac, frac) per axis; nodes on the box boundary move tangentially only."""
    n = mesh.nnodes
    idx = np.arange(n, dtype=np.uint64)
    dx = h * (2 * frac * uniform(seed, stream, 2 * idx) - frac)
    dy = h * (2 * frac * uniform(seed, stream, 2 * idx + 1) - frac)
    x, y = mesh.xy[:, 0], mesh.xy[:, 1]
    tol = 1e-12 * max(1.0, box[1] - box[0])
    onx = (np.abs(x - box[0]) < tol) | (np.abs(x - box[1]) < tol)
    ony = (np.abs(y - box[2]) < tol) | (np.abs(y - box[3]) < tol)
    dx[onx] = 0.0
    dy[ony] = 0.0
    mesh.xy = np.ascontiguousarray(np.stack([x + dx, y + dy], 1))
    return mesh


def heat_t3(N=100):
    mesh = T3block(1.0, 1.0, N, N)
    fesp = FESpace(mesh, FEH1_T3())
    b = mesh.boundary_nodes()
    xy = mesh.xy[b - 1]
    fesp.setebc(0, b, 1, 1.0 + xy[:, 0] ** 2 + 2.0 * xy[:, 1] ** 2)
    numberdofs(fesp)
    return dict(name=f"heat-T3 {N}x{N} 1-pt", mesh=mesh, fesp=fesp, kappa=1.0, Q=-6.0,
                qsettings=dict(kind="default", npts=1), ndofs=fesp.ndofs(), nfree=fesp.nunknowns())


def elast_t6(N=400, frac=0.15, seed=SEED, E=1.0, nu=0.3):
    mesh = T6block(1.0, 1.0, N, N)
    if frac > 0:
        jitter(mesh, 1.0 / N, frac, seed)
    fesp = FESpace(mesh, FEH1_T6(), 2)
    x = mesh.xy[:, 0]
    left = np.nonzero(np.abs(x) < 1e-12)[0] + 1
    right = np.nonzero(np.abs(x - 1.0) < 1e-12)[0] + 1
    fesp.setebc(0, left, 1, 0.0)
    fesp.setebc(0, left, 2, 0.0)
    fesp.setebc(0, right, 1, 0.1)
    fesp.setebc(0, right, 2, 0.0)
    numberdofs(fesp)
    return dict(name=f"elasticity-T6 {N}x{N} 3-pt plane strain", mesh=mesh, fesp=fesp, D=plane_strain_D(E, nu),
                qsettings=dict(kind="default", npts=3), ndofs=fesp.ndofs(), nfree=fesp.nunknowns())


def stokes_t6t3(N=500, mu=1.0):
    vmesh = T6block(1.0, 1.0, N, N)
    pmesh = T6toT3(vmesh)
    u = FESpace(vmesh, FEH1_T6(), 2)
    b = vmesh.boundary_nodes()
    u.setebc(0, b, 1, 0.0)
    u.setebc(0, b, 2, 0.0)
    p = FESpace(pmesh, FEH1_T3(), 1)
    p.setebc(0, pmesh.nearestto([0.5, 0.5]), 1, 0.0)
    numberdofs([u, p])
    return dict(name=f"Stokes T6/T3 {N}x{N} 3-pt", vmesh=vmesh, pmesh=pmesh, ufesp=u, pfesp=p, mu=mu,
                D=np.diag([2 * mu, 2 * mu, mu]), qsettings=dict(kind="default", npts=3),
                ndofs=u.ndofs() + p.ndofs(), nfree=u.nunknowns() + p.nunknowns())


def heat_q4(N=4096, sigma=1.0, seed=SEED):
    mesh = Q4block(1.0, 1.0, N, N)
    fesp = FESpace(mesh, FEH1_Q4())
    b = mesh.boundary_nodes()
    fesp.setebc(0, b, 1, 0.0)
    numberdofs(fesp)
    E = mesh.nelements
    kappa_e = np.exp(sigma * normal(seed, 3, E))
    Q_e = 2.0 * uniform(seed, 4, np.arange(E, dtype=np.uint64)) - 1.0
    return dict(name=f"heat-Q4 {N}x{N} 2x2 Gauss", mesh=mesh, fesp=fesp, kappa_e=kappa_e, Q_e=Q_e,
                qsettings=dict(kind="Gauss", order=2), ndofs=fesp.ndofs(), nfree=fesp.nunknowns())


def batch_specs(nprob=256, seed=SEED, emin=1e3, emax=2e5):
    """Kinds and sizes of the batched problems: heat-T3 if U < 0.5 else
    elasticity-T6; E_k log-uniform in [emin, emax], rounded to a square grid
    (E = 2 n^2)."""
    u_kind = uniform(seed, 5, np.arange(nprob, dtype=np.uint64))
    u_size = uniform(seed, 6, np.arange(nprob, dtype=np.uint64))
    Et = np.exp(np.log(emin) + u_size * (np.log(emax) - np.log(emin)))
    n = np.maximum(2, np.round(np.sqrt(Et / 2))).astype(int)
    return [("heat" if u_kind[k] < 0.5 else "elast", int(n[k])) for k in range(nprob)]


def batch_problem(k, kind, n, seed=SEED, sigma=0.5, frac=0.20):
    """Problem k of the batch: jittered mesh and lognormal per-element coefficient."""
    if kind == "heat":
        mesh = T3block(1.0, 1.0, n, n)
        jitter(mesh, 1.0 / n, frac, seed, stream=100 + 3 * k)
        fesp = FESpace(mesh, FEH1_T3())
        b = mesh.boundary_nodes()
        fesp.setebc(0, b, 1, 0.0)
        numberdofs(fesp)
        coef = np.exp(sigma * normal(seed, 101 + 3 * k, mesh.nelements))
        return dict(kind="heat", mesh=mesh, fesp=fesp, kappa_e=coef, Q=1.0, qsettings=dict(kind="default", npts=1),
                    ndofs=fesp.ndofs())
    mesh = T6block(1.0, 1.0, n, n)
    jitter(mesh, 1.0 / n, frac, seed, stream=100 + 3 * k)
    fesp = FESpace(mesh, FEH1_T6(), 2)
    x = mesh.xy[:, 0]
    left = np.nonzero(np.abs(x) < 1e-12)[0] + 1
    fesp.setebc(0, left, 1, 0.0)
    fesp.setebc(0, left, 2, 0.0)
    numberdofs(fesp)
    coef = np.exp(sigma * normal(seed, 102 + 3 * k, mesh.nelements))
    return dict(kind="elast", mesh=mesh, fesp=fesp, scale_e=coef, D=plane_strain_D(1.0, 0.3),
                qsettings=dict(kind="default", npts=3), ndofs=fesp.ndofs())
