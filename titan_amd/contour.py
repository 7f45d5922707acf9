"""Energy contour and k-meshes (host set-up of the (k,E) work list).

  gauleg / imag_epoints / real_epoints   EnergyIntegration.F90:45-157
  omega grid                            setLoops.F90:14-19, mod_io.F90:482
  uniform k mesh                        mod_BrillouinZone.F90:151-472 (the
      reference folds the same points into the first zone and splits
      boundary duplicates; for k-periodic H the sums are identical)
  adaptive imaginary-axis mesh sizes    adaptiveMesh.F90:23-63, 199-218
"""
import math

import numpy as np


def gauleg(x1, x2, n):
    """Gauss-Legendre nodes/weights on [x1, x2]: Newton on P_n from the
    Tricomi start cos(pi (i - 1/4)/(n + 1/2)), stop at |dz| <= 3e-14."""
    m = (n + 1) // 2
    xm, xl = 0.5 * (x2 + x1), 0.5 * (x2 - x1)
    x = np.zeros(n)
    w = np.zeros(n)
    for i in range(1, m + 1):
        z = math.cos(3.141592654 * (i - 0.25) / (n + 0.5))
        while True:
            p1, p2 = 1.0, 0.0
            for j in range(1, n + 1):
                p3, p2 = p2, p1
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j
            pp = n * (z * p1 - p2) / (z * z - 1.0)
            z1 = z
            z = z1 - p1 / pp
            if abs(z - z1) <= 3e-14:
                break
        x[i - 1] = xm - xl * z
        x[n - i] = xm + xl * z
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp)
        w[n - i] = w[i - 1]
    return x, w


def imag_epoints(parts, n1gl):
    """y nodes (ascending) and weights of int_0^inf dy via y = x/(1-x)."""
    xs, ps = [], []
    for k in range(parts, 0, -1):
        if parts != 1:
            if k == 1:
                e1, e2 = 1e-2, 1.0
            elif k != parts:
                e1, e2 = 10.0 ** (-k - 1), 10.0 ** (-k)
            else:
                e1, e2 = 0.0, 10.0 ** (-k)
        else:
            e1, e2 = 0.0, 1.0
        x, p = gauleg(e1, e2, n1gl)
        xs.append((k, x))
        ps.append((k, p))
    x1 = np.concatenate([x for _, x in sorted(xs)])
    p1 = np.concatenate([p for _, p in sorted(ps)])
    y = x1 / (1.0 - x1)
    wy = p1 / (1.0 - x1) ** 2
    order = np.argsort(y, kind="stable")
    return y[order], wy[order]


def real_epoints(Ef, e, parts3, n3gl):
    """Nodes/weights on [Ef - e, Ef] in parts3 equal segments."""
    xs, ps = [], []
    for k in range(1, parts3 + 1):
        e1 = Ef - ((parts3 - (k - 1)) * e / parts3)
        e2 = Ef - ((parts3 - k) * e / parts3)
        x, p = gauleg(e1, e2, n3gl)
        xs.append(x)
        ps.append(p)
    return np.concatenate(xs), np.concatenate(ps)


def omega_grid(emin, emax, nEner):
    """emin + (k-1) (emax-emin)/nEner, k = 1..nEner+1 (a single point when
    the step is below 1e-15)."""
    if nEner < 1 or (emax - emin) / max(nEner, 1) <= 1e-15:
        return np.array([emin])
    de = (emax - emin) / nEner
    return emin + de * np.arange(nEner + 1)


def uniform_kmesh(b, n1, n2=1, n3=1):
    """k = sum_i n_i b_i / N_i with weights 1/(N1 N2 N3)."""
    b = np.asarray(b, dtype=float)
    i, j, k = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    kp = (i.reshape(-1, 1) * b[0] / n1 + j.reshape(-1, 1) * b[1] / n2 + k.reshape(-1, 1) * b[2] / n3)
    w = np.full(len(kp), 1.0 / (n1 * n2 * n3))
    return kp, w


def mesh_dims(nkpt, sysdim):
    """Per-dimension points of an nkpt mesh (mod_io.F90:103-121)."""
    if sysdim == 3:
        n = math.ceil(float(nkpt) ** 0.333333333333333)
        return (n, n, n)
    if sysdim == 2:
        n = math.ceil(float(nkpt) ** 0.5)
        return (n, n, 1)
    return (int(nkpt), 1, 1)


def adaptive_nkpt(y, y1, nkpt_total, sysdim, minimum):
    """N(y) = max(ceil(N / (y/y1)^sqrt(d)), minimumBZmesh) (adaptiveMesh.F90:199-218)."""
    n = math.ceil(nkpt_total / ((y / y1) ** math.sqrt(sysdim)))
    return max(n, int(minimum))


def imag_work_list(model, y, wy, nkpt, minimum=None):
    """Flattened (E,k) list of the imaginary axis: k[n,3], y[n], w[n]
    (w = wght(E) w_k), one mesh per energy (adaptive when minimum < nkpt)."""
    sd = model.sysdim
    minimum = nkpt if minimum is None else min(minimum, nkpt)
    ks, ys, ws = [], [], []
    cache = {}
    for e in range(len(y)):
        n = adaptive_nkpt(y[e], y[0], nkpt, sd, minimum)
        dims = mesh_dims(n, sd)
        if dims not in cache:
            cache[dims] = uniform_kmesh(model.b, *dims)
        kp, wk = cache[dims]
        ks.append(kp)
        ys.append(np.full(len(wk), y[e]))
        ws.append(wy[e] * wk)
    return np.concatenate(ks), np.concatenate(ys), np.concatenate(ws)
