"""Gradients of chi(q, omega) with respect to the physical parameters of the
tight-binding model, through torch.autograd.

The reference computes chi but no derivatives.  Here the GPU forward
(th_chi0 + th_dyson) and its adjoint (th_chi_adjoint, which recomputes the
Green functions instead of storing them) sit behind one
torch.autograd.Function, ChiFunction, called over ctypes; the map from the
parameters to the tables the library takes (hops t, local Hamiltonian h0,
interaction matrix U) is small dense torch code in TorchModel, so
torch.autograd carries the chain rule to

  sk      [nTypes, nStages, 10]  two-centre Slater-Koster integrals
                                 (TightBinding.F90, mixed across types by the
                                 geometric-mean rule of the reference; that
                                 rule has a kink where an integral of one of
                                 two different types is zero, and the
                                 gradient there is taken as zero)
  lam     [nTypes, 2]            SOC strengths lambda_p, lambda_d (mod_SOC.F90)
  U       [nTypes, 2]            Un, Um: the interaction matrix
                                 (mod_susceptibilities.F90:106-159) and the
                                 mean-field term hee (mod_U_matrix.F90:43-52)
                                 at fixed ground-state occupations and moments
  onsite  [nTypes, 9]            on-site energies by orbital id

Usage:

    tm = TorchModel(spec)
    p = tm.params()                       # leaf tensors, requires_grad=True
    op = ChiOperator(tm, kmesh=(8, 8, 8), n1gl=32, n3gl=16, eta=5e-3)
    chi = op(p, omega)                    # complex128 D x D, differentiable
    L = some_loss(op.trace(chi)); L.backward(); p["U"].grad ...
"""
import numpy as np
import torch

from titan_amd import binding as B
from titan_amd import contour as Ct
from titan_amd.system import TBModel, sk_table


class TorchModel:
    """Differentiable tables (t, h0, U) of one system as functions of its
    physical parameters; the geometry (neighbour shells, direction cosines,
    orbital selection, occupations) is fixed by the spec."""

    def __init__(self, spec):
        self.spec = spec
        self.m = m = TBModel(spec)
        nO, nA = m.nOrb, m.nA
        o = m.orbs - 1
        nhop = len(m.R)
        E = np.zeros((nhop, 10, nO, nO))
        for h in range(nhop):
            for l in range(10):
                e = np.zeros(10)
                e[l] = 1.0
                E[h, l] = sk_table(e, m.dircos[h])[np.ix_(o, o)]
        self.E = torch.from_numpy(E)
        self.type_i = torch.from_numpy(m.type[m.ib].astype(np.int64))     # 'i' end of each hop (ib)
        self.type_j = torch.from_numpy(m.type[m.jb].astype(np.int64))     # 'j' end (jb)
        self.stage = torch.from_numpy(m.stage_of.astype(np.int64))
        scale = float(spec.get("socscale", 1.0))
        self.Lp = torch.from_numpy(m.soc_block(1.0, 0.0) * scale)
        self.Ld = torch.from_numpy(m.soc_block(0.0, 1.0) * scale)
        self.Xn = torch.from_numpy(np.stack([m.exchange_block(i, 1.0, 0.0) for i in range(nA)]))
        self.Xm = torch.from_numpy(np.stack([m.exchange_block(i, 0.0, 1.0) for i in range(nA)]))
        D = m.D
        idx, cn, cm, at = [], [], [], []
        for i in range(nA):
            on_n = [(1.0, 0.0) if a == i else (0.0, 0.0) for a in range(nA)]
            on_m = [(0.0, 1.0) if a == i else (0.0, 0.0) for a in range(nA)]
            Cn = m.interaction_matrix(on_n).ravel()
            Cm = m.interaction_matrix(on_m).ravel()
            nz = np.nonzero((Cn != 0) | (Cm != 0))[0]
            idx.append(nz)
            cn.append(Cn[nz].real)
            cm.append(Cm[nz].real)
            at.append(np.full(len(nz), i))
        self.u_idx = torch.from_numpy(np.concatenate(idx).astype(np.int64))
        self.u_cn = torch.from_numpy(np.concatenate(cn))
        self.u_cm = torch.from_numpy(np.concatenate(cm))
        self.u_at = torch.from_numpy(np.concatenate(at).astype(np.int64))
        self.atom_type = torch.from_numpy(m.type.astype(np.int64))

    def params(self, requires_grad=True):
        ty = self.spec["types"]
        nS = self.m.nStages

        def leaf(a):
            return torch.tensor(np.asarray(a, dtype=np.float64), requires_grad=requires_grad)
        return {"sk": leaf([np.asarray(t["sk"], dtype=float)[:nS] for t in ty]),
                "lam": leaf([t["lambda"] for t in ty]),
                "U": leaf([t["U"] for t in ty]),
                "onsite": leaf([t["onsite"] for t in ty])}

    def hops(self, p):
        """t [nhop, nOrb, nOrb] (real; rows = orbitals of jb)."""
        t1 = p["sk"][self.type_i, self.stage]          # [nhop, 10]
        t2 = p["sk"][self.type_j, self.stage]
        same = (self.type_i == self.type_j)[:, None]
        prod = torch.abs(t1 * t2)
        safe = torch.where(prod > 0, prod, torch.ones_like(prod))
        g = torch.where(prod > 0, torch.sqrt(safe), torch.zeros_like(prod))
        mixed = torch.where(t1 + t2 >= 0, g, -g)
        mix = torch.where(same, t1, mixed)
        return torch.einsum("hl,hlab->hab", mix, self.E)

    def local(self, p):
        """h0 [d, d] complex: on-site + SOC + mean-field exchange per atom."""
        m = self.m
        nO, nA, d = m.nOrb, m.nA, m.d
        o = torch.from_numpy((m.orbs - 1).astype(np.int64))
        blocks = []
        for i in range(nA):
            ty = int(m.type[i])
            on = p["onsite"][ty][o]
            dg = torch.diag(torch.cat([on, on])).to(torch.complex128)
            lam = p["lam"][ty]
            Un, Um = p["U"][ty][0], p["U"][ty][1]
            blocks.append(dg + lam[0] * self.Lp + lam[1] * self.Ld + Un * self.Xn[i] + Um * self.Xm[i])
        return torch.block_diag(*blocks) if nA > 1 else blocks[0]

    def umat(self, p):
        """Interaction matrix [D, D] complex (chi index basis)."""
        D = self.m.D
        ty = self.atom_type[self.u_at]
        vals = p["U"][ty, 0] * self.u_cn + p["U"][ty, 1] * self.u_cm
        flat = torch.zeros(D * D, dtype=torch.float64).index_add(0, self.u_idx, vals)
        return flat.reshape(D, D).to(torch.complex128)

    def tables(self, p):
        return self.hops(p), self.local(p), self.umat(p)


class ChiFunction(torch.autograd.Function):
    """chi = Dyson(chi0(t, h0), U) on the GPU; backward = th_chi_adjoint."""

    @staticmethod
    def forward(ctx, t, h0, U, op, omega):
        h = op.handle(t.detach().cpu().numpy(), h0.detach().cpu().numpy())
        x2, p2 = op.real_nodes(omega)
        chi0 = h.chi0(op.q, omega, op.Ef, op.eta, op.etap, x2, p2, precision=op.precision)
        Un = U.detach().cpu().numpy()
        chi, cond = h.dyson(chi0, Un)
        ctx.h, ctx.op, ctx.omega = h, op, omega
        ctx.chi0, ctx.chi, ctx.U = chi0, chi, Un
        ctx.t_real = not t.is_complex()
        ctx.cond = cond
        return torch.from_numpy(np.array(chi))

    @staticmethod
    def backward(ctx, g):
        op = ctx.op
        x2, p2 = op.real_nodes(ctx.omega)
        gh = g.detach().resolve_conj().cpu().numpy()
        th, hh, uh = ctx.h.adjoint(op.q, ctx.omega, op.Ef, op.eta, op.etap, ctx.chi0, gh, U=ctx.U, chi=ctx.chi,
                                   x2=x2, p2=p2, precision=op.adjoint_precision)
        gt = torch.from_numpy(th.real.copy() if ctx.t_real else th)
        return gt, torch.from_numpy(hh), torch.from_numpy(uh), None, None


class ChiOperator:
    """One system on one k-mesh and contour; op(p, omega) is the
    differentiable chi(q, omega) of parameters p (TorchModel.params())."""

    def __init__(self, tm, kmesh, n1gl=64, parts=1, n3gl=32, parts3=1, eta=5e-4, etap=None, q=(0.0, 0.0, 0.0),
                 precision=B.PREC_FP64, adjoint_precision=None, tuning=None):
        self.tm = tm
        self.tuning = dict(tuning or {})   # th_tune settings of every handle (e.g. an autotuner's best)
        m = tm.m
        self.Ef = m.Ef
        self.eta = float(eta)
        self.etap = float(eta if etap is None else etap)
        self.q = np.asarray(q, dtype=float)
        self.parts3, self.n3gl = parts3, n3gl
        self.precision = precision
        self.adjoint_precision = precision if adjoint_precision is None else adjoint_precision
        self._h, self._key = None, None
        y, wy = Ct.imag_epoints(parts, n1gl)
        self.kp, self.wk = Ct.uniform_kmesh(m.b, *kmesh)
        self.kim = np.tile(self.kp, (len(y), 1))
        self.yim = np.repeat(y, len(self.wk))
        self.wim = np.outer(wy, self.wk).ravel()
        nO, nA, D = m.nOrb, m.nA, m.D
        # orbital trace schi[(sigma,i),(sigma',j)] = sum chi[(sigma,i,mu,mu),(sigma',j,nu,nu)]
        S = np.zeros((4 * nA, D))
        for sg in range(4):
            for i in range(nA):
                for mu in range(nO):
                    S[sg * nA + i, m.cidx(sg, i, mu, mu)] = 1.0
        self.S = torch.from_numpy(S).to(torch.complex128)

    def real_nodes(self, omega):
        if abs(omega) < 1e-15:
            return None, None
        return Ct.real_epoints(self.Ef, omega, self.parts3, self.n3gl)

    def handle(self, t, h0):
        """A handle of the model (t, h0): the one of the previous call while the
        tables are unchanged, so that the frequencies of one loss share its
        eigen-decomposition and omega-independent panels (a handle a pending
        backward still holds stays alive through that reference)."""
        m = self.tm.m
        t = np.asarray(t, dtype=complex)
        h0 = np.asarray(h0, dtype=complex)
        key = (t.tobytes(), h0.tobytes())
        if self._h is not None and self._key == key:
            return self._h
        h = B.Handle(m.nA, m.nOrb, m.R, m.jb, m.ib, t, h0)
        if self.tuning:
            h.tune(**self.tuning)
        h.set_imag_points(self.kim, self.yim, self.wim)
        h.set_real_kmesh(self.kp, self.wk)
        self._h, self._key = h, key
        return h

    def close(self):
        """Frees the kept handle's device memory (a later call builds a new one)."""
        if self._h is not None:
            self._h.close()
        self._h, self._key = None, None

    def __call__(self, p, omega):
        t, h0, U = self.tm.tables(p)
        return ChiFunction.apply(t, h0, U, self, float(omega))

    def trace(self, chi):
        """schi = S chi S^T (4nA x 4nA, index sigma*nA + i)."""
        return self.S @ chi @ self.S.T

    def chi_pm(self, chi):
        """The transverse susceptibility chi^{+-} summed over the sites: the
        (+, +) block of schi (sigma = + for both indices in the reference's
        labelling, its chi_i_j files), sum_ij schi[(+, i), (+, j)]."""
        nA = self.tm.m.nA
        s = self.trace(chi)
        return torch.sum(s[0:nA, 0:nA])


def target_curve(omegas, scale, seed=3):
    """A seeded target for Im chi^{+-}(omega) of magnitude 'scale' (callers
    take the largest |Im chi^{+-}| of the model at its nominal parameters): a
    Lorentzian of seeded position, width and sign over the frequency window
    plus a seeded linear background, zero at omega = 0."""
    rng = np.random.default_rng(seed)
    om = np.asarray(omegas, dtype=float)
    top = max(float(om.max()), 1e-12)
    w0 = top * rng.uniform(0.2, 0.8)
    gam = top * rng.uniform(0.05, 0.2)
    sign = -1.0 if rng.uniform() < 0.5 else 1.0
    slope = rng.uniform(-0.3, 0.3)
    lor = gam * om / ((om - w0) ** 2 + gam ** 2)
    lor = lor / max(float(np.max(np.abs(lor))), 1e-300)
    return float(scale) * (sign * lor + slope * om / top)


def nominal_chi_pm(op, tm, omegas):
    """Im chi^{+-}(omega) of the model at its nominal parameters (no graph)."""
    with torch.no_grad():
        p = tm.params(requires_grad=False)
        return np.array([float(torch.imag(op.chi_pm(op(p, float(om))))) for om in omegas])


def chi_pm_loss(op, p, omegas, target):
    """The adjoint workload's scalar loss: sum over omega of
    (Im chi^{+-}(omega) - target(omega))^2, differentiable in p."""
    L = torch.zeros((), dtype=torch.float64)
    for om, tg in zip(omegas, target):
        L = L + (torch.imag(op.chi_pm(op(p, float(om)))) - float(tg)) ** 2
    return L
