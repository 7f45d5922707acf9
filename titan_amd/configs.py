"""Seeded synthetic systems of the benchmark and parity workloads.

Energies are drawn in eV and converted to Ry with TITAN's ry2ev = 13.6
(mod_io.F90:183).  Every system is a plain dict accepted by
titan_amd.system.TBModel; run parameters (mesh, contour, omega grid) come
with it.  The seed is stored in the dict.

  cfg0  bcc, 1 site, s+p+d (18 x 18 H), 2 shells, 16^3 k, 64 + 32 energies,
        16 omega in [0, 0.01] Ry, eta = 5e-4 Ry
  cfg1  fcc(001) slab, 8 sites (144 x 144 H), layer-dependent on-site
        shifts, 128^2 k, 64 + 32 energies, 32 omega
  cfg2  bcc supercell, 3 sites (54 x 54 H) of the cfg0 element with U_d =
        3 eV, 48^3 k, 64 + 32 energies, 64 omega clustered at the
        transverse (FMR) pole of its self-consistent ground state
        (omega_grid)
  cfg3  cfg0 system at 24^3 k, 32 omega (adjoint workload)

A seeded moment is not self-consistent, so the RPA susceptibility of a
seeded state has no Goldstone/FMR pole (cfg0 keeps its seeded m = 2.2 muB
at U_d = 1 eV, as specified).  With U_d = 1 eV the seeded element's
self-consistent state is nonmagnetic (checked on the GPU at 48^3 k-points:
m = 0, Dyson condition numbers below 10 from -0.3 to 0.5 Ry), so cfg2, whose
frequencies must sit at the pole, takes U_d = 3 eV, where the element is a
self-consistent ferromagnet (3.85 muB of d moment per atom at cfg2's mesh).  sc_states.json (next
to this module) holds the self-consistent d moments and Fermi level that
titan_amd.selfconsistency computes on the GPU for cfg2's own k-mesh and
contour (python -m titan_amd.selfconsistency --config 2), and the pole it
then finds; system(cfg) applies them where present.
"""
import json
import os

import numpy as np

_SC_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sc_states.json")

RY2EV = 13.6


def _ry(x):
    return np.asarray(x, dtype=float) / RY2EV


def seeded_type(rng, nstages):
    """On-site energies and SK integrals of one element, in Ry."""
    sk = []
    for _ in range(nstages):
        p = rng.uniform(-1.2, 1.2, size=10)
        p[0] = rng.uniform(-1.5, -0.5)          # sss
        sk.append(_ry(p).tolist())
    on = rng.uniform(-2.0, 8.0, size=4)           # s, p, t2g, eg
    onsite = _ry([on[0], on[1], on[1], on[1], on[2], on[2], on[2], on[3], on[3]]).tolist()
    return {"onsite": onsite, "sk": sk}


def _bcc_vectors(a0):
    return (0.5 * a0 * np.array([[1, 1, -1], [1, -1, 1], [-1, 1, 1]], dtype=float))


def fermi_level(model, filling, nk=8):
    """E_F at which the k-averaged occupation of H(k) equals 'filling'
    electrons per cell (bisection on a uniform nk^sysdim mesh)."""
    from titan_amd.contour import uniform_kmesh
    dims = (nk, nk, nk) if model.sysdim == 3 else ((nk, nk, 1) if model.sysdim == 2 else (nk, 1, 1))
    kp, w = uniform_kmesh(model.b, *dims)
    ev = np.array([np.linalg.eigvalsh(model.hk(k)) for k in kp])
    lo, hi = ev.min(), ev.max()
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        n = np.sum(w[:, None] * (ev <= mid))
        if n < filling:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def _finish(spec, filling):
    from titan_amd.system import TBModel
    spec["Ef"] = 0.0
    m = TBModel(spec)
    spec["Ef"] = float(fermi_level(m, filling))
    return spec


def bcc_1site(seed=0, a0=5.30, nstages=2, lam_d_ev=0.06, U_ev=1.0, m_d=2.2):
    rng = np.random.default_rng(seed)
    ty = seeded_type(rng, nstages)
    ty["lambda"] = [0.0, float(_ry(lam_d_ev))]
    ty["U"] = [0.0, float(_ry(U_ev))]
    mz = [0.0] * 4 + [m_d / 5] * 5
    spec = {
        "name": "bcc 1-site seeded", "seed": seed, "sysdim": 3, "a": _bcc_vectors(a0).tolist(),
        "pos": [[0.0, 0.0, 0.0]], "type": [0], "orbs": list(range(1, 10)), "nStages": nstages,
        "relTol": 0.05, "types": [ty], "rho": [[0.0] * 9], "mx": [[0.0] * 9], "my": [[0.0] * 9],
        "mz": [mz], "socscale": 1.0,
    }
    return _finish(spec, filling=8.0)


def bcc_supercell(nsites=3, seed=2, a0=5.30, nstages=2, lam_d_ev=0.06, U_ev=1.0, m_d=2.2):
    """nsites bcc primitive cells stacked along a1, one element."""
    rng = np.random.default_rng(seed)
    ty = seeded_type(rng, nstages)
    ty["lambda"] = [0.0, float(_ry(lam_d_ev))]
    ty["U"] = [0.0, float(_ry(U_ev))]
    a = _bcc_vectors(a0)
    A = np.array([nsites * a[0], a[1], a[2]])
    pos = [(k * a[0]).tolist() for k in range(nsites)]
    mz = [0.0] * 4 + [m_d / 5] * 5
    spec = {
        "name": "bcc %d-site supercell seeded" % nsites, "seed": seed, "sysdim": 3, "a": A.tolist(),
        "pos": pos, "type": [0] * nsites, "orbs": list(range(1, 10)), "nStages": nstages, "relTol": 0.05,
        "types": [ty], "rho": [[0.0] * 9] * nsites, "mx": [[0.0] * 9] * nsites, "my": [[0.0] * 9] * nsites,
        "mz": [mz] * nsites, "socscale": 1.0,
    }
    return _finish(spec, filling=8.0 * nsites)


def fcc001_slab(nlayers=8, seed=1, a0=6.65, nstages=2, lam_d_ev=0.06, U_ev=1.0, m_d=2.2):
    """fcc(001) slab: square in-plane lattice a0/sqrt(2), layers a0/2 apart,
    ABAB stacking; one element with layer-dependent on-site shifts."""
    rng = np.random.default_rng(seed)
    base = seeded_type(rng, nstages)
    types = []
    for _ in range(nlayers):
        t = {"onsite": (np.asarray(base["onsite"]) + _ry(rng.uniform(-0.3, 0.3))).tolist(), "sk": base["sk"],
             "lambda": [0.0, float(_ry(lam_d_ev))], "U": [0.0, float(_ry(U_ev))]}
        types.append(t)
    h = 0.5 * a0
    a = np.array([[h, h, 0.0], [-h, h, 0.0], [0.0, 0.0, 0.0]])
    pos = [[(l % 2) * h, 0.0, l * h] for l in range(nlayers)]
    mz = [0.0] * 4 + [m_d / 5] * 5
    spec = {
        "name": "fcc(001) %d-layer slab seeded" % nlayers, "seed": seed, "sysdim": 2, "a": a.tolist(),
        "pos": pos, "type": list(range(nlayers)), "orbs": list(range(1, 10)), "nStages": nstages,
        "relTol": 0.05, "types": types, "rho": [[0.0] * 9] * nlayers, "mx": [[0.0] * 9] * nlayers,
        "my": [[0.0] * 9] * nlayers, "mz": [mz] * nlayers, "socscale": 1.0,
    }
    return _finish(spec, filling=8.0 * nlayers)


def run_params(cfg):
    """Mesh / contour / omega parameters of each configuration."""
    if cfg == 0:
        return {"nk": (16, 16, 16), "parts": 1, "n1gl": 64, "parts3": 1, "n3gl": 32, "eta": 5e-4, "etap": 5e-4,
                "omega": np.linspace(0.0, 0.01, 16).tolist(), "q": [0.0, 0.0, 0.0]}
    if cfg == 1:
        return {"nk": (128, 128, 1), "parts": 1, "n1gl": 64, "parts3": 1, "n3gl": 32, "eta": 5e-4, "etap": 5e-4,
                "omega": np.linspace(0.0, 0.01, 32).tolist(), "q": [0.0, 0.0, 0.0]}
    if cfg == 2:
        return {"nk": (48, 48, 48), "parts": 1, "n1gl": 64, "parts3": 1, "n3gl": 32, "eta": 5e-4, "etap": 5e-4,
                "omega": None, "q": [0.0, 0.0, 0.0]}
    if cfg == 3:
        return {"nk": (24, 24, 24), "parts": 1, "n1gl": 64, "parts3": 1, "n3gl": 32, "eta": 5e-4, "etap": 5e-4,
                "omega": np.linspace(0.0, 0.01, 32).tolist(), "q": [0.0, 0.0, 0.0]}
    raise ValueError("unknown configuration %r" % cfg)


def uneven_cell_sizes(ncells=40, seed=4):
    """Atoms per cell of the uneven batch, drawn from {1, 2, 4, 8, 16}."""
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.choice([1, 2, 4, 8, 16], size=ncells)]


def uneven_kmesh(nA):
    """Per-cell k-mesh edge of the uneven batch: 24 at nA = 1 down to 8 at
    nA = 16, scaled as nA^(-log 3 / log 16) (inversely with the cell size)."""
    return int(round(24.0 / nA ** (np.log(3.0) / np.log(16.0))))


def uneven_batch(ncells=40, seed=4, sizes=None):
    """The uneven batch: cells of n_atoms in {1, 2, 4, 8, 16} (H from 18 x 18
    to 288 x 288), each a bcc supercell with its own seeded Slater-Koster,
    SOC and U parameters as in the 1-site system, per-cell k-mesh from 24^3
    (nA = 1) to 8^3 (nA = 16), 48 imaginary + 16 real energies (64 per k at
    omega != 0), 16 frequencies in (0, 0.01] Ry.  Returns [(spec, run)]."""
    sizes = uneven_cell_sizes(ncells, seed) if sizes is None else list(sizes)
    cells = []
    for c, nA in enumerate(sizes):
        spec = bcc_supercell(nA, seed=1000 + seed * 100 + c)
        spec["name"] = "uneven cell %d: bcc %d-site" % (c, nA)
        n = uneven_kmesh(nA)
        run = {"nk": (n, n, n), "parts": 1, "n1gl": 48, "parts3": 1, "n3gl": 16, "eta": 5e-4, "etap": 5e-4,
               "omega": np.linspace(0.01 / 16, 0.01, 16).tolist(), "q": [0.0, 0.0, 0.0]}
        cells.append((spec, run))
    return cells


def seeded_system(cfg):
    """The seeded system of a configuration (moments as drawn, E_F at the
    filling on an 8^3 mesh)."""
    if cfg in (0, 3):
        return bcc_1site()
    if cfg == 1:
        return fcc001_slab()
    if cfg == 2:
        return bcc_supercell(3, seed=0, U_ev=3.0)
    raise ValueError("unknown configuration %r" % cfg)


def filling(cfg):
    """Electrons per cell of a configuration (8 per site)."""
    return 8.0 * {0: 1, 1: 8, 2: 3, 3: 1}[cfg]


def sc_state(cfg):
    """Self-consistent state of a configuration from sc_states.json, or None."""
    key = str(cfg)
    if not os.path.exists(_SC_PATH):
        return None
    with open(_SC_PATH) as fh:
        return json.load(fh).get(key)


def system(cfg):
    """The configuration's system: seeded parameters with the self-consistent
    d moments and Fermi level of sc_states.json where present."""
    spec = seeded_system(cfg)
    st = sc_state(cfg)
    if st is None:
        return spec
    dorb = [k for k, o in enumerate(spec["orbs"]) if o >= 5]
    mz = np.asarray(spec["mz"], dtype=float).copy()
    for i, m in enumerate(st["mz_d"]):
        mz[i, :] = 0.0
        mz[i, dorb] = m / len(dorb)
    spec["mz"] = mz.tolist()
    spec["Ef"] = float(st["Ef"])
    spec["name"] += " (self-consistent)"
    return spec


def clustered_omegas(center, half_width, n=64, span=16.0, omax=0.01):
    """n frequencies in (0, omax] clustered at a resonance: the quantiles
    (k + 1/2)/n of a Lorentzian of the given centre and half width, cut to
    [center - span*half_width, center + span*half_width] and to (0, omax]."""
    lo = max(center - span * half_width, 1e-3 * half_width)
    hi = min(center + span * half_width, omax)
    a, b = np.arctan((lo - center) / half_width), np.arctan((hi - center) / half_width)
    u = (np.arange(n) + 0.5) / n
    return center + half_width * np.tan(a + u * (b - a))


def omega_grid(cfg, spec=None):
    """Frequencies of a configuration (Ry).  cfg2: 64 frequencies clustered
    at the transverse resonance of its self-consistent state
    (clustered_omegas at the pole and half width of sc_states.json: the
    maximum of |Im chi^{+-}| found by titan_amd.selfconsistency.locate_pole
    with FP64 sweeps); without a stored state, 64 uniform frequencies up to
    0.01 Ry."""
    r = run_params(cfg)
    if r["omega"] is not None:
        return np.asarray(r["omega"], dtype=float)
    st = sc_state(cfg)
    if st is not None and "omega_pole" in st:
        return clustered_omegas(st["omega_pole"], st["half_width"])
    return np.linspace(0.01 / 64, 0.01, 64)


def chain_p(nA=2, seed=5, a0=4.0):
    """Small 1D chain with p_x, p_y orbitals (no d orbitals: U = 0), for
    checks of other orbital counts and of the 1D k-mesh."""
    rng = np.random.default_rng(seed)
    types = []
    for i in range(nA):
        ty = seeded_type(rng, 1)
        ty["lambda"] = [float(_ry(0.2)), 0.0]
        ty["U"] = [0.0, 0.0]
        types.append(ty)
    pos = [[0.0, (a0 / nA) * i * 0.5, 0.0] for i in range(nA)]
    spec = {
        "name": "p chain seeded", "seed": seed, "sysdim": 1, "a": [[a0, 0, 0], [0, 0, 0], [0, 0, 0]],
        "pos": pos, "type": list(range(nA)), "orbs": [2, 3], "nStages": 1, "relTol": 0.05, "types": types,
        "rho": [[0.0] * 2] * nA, "mx": [[0.0] * 2] * nA, "my": [[0.0] * 2] * nA, "mz": [[0.0] * 2] * nA,
        "socscale": 1.0,
    }
    return _finish(spec, filling=1.0 * nA)
